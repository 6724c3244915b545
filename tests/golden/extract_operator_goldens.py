"""Extract the operator goldens (y.y values) that pTatin3D's own integration tests
hold in tests/integration/operator/**/expected, together with the mesh and model
options of the matching opts file, into operator_goldens.json (data only).

Run in a container where the reference tree is mounted:
    python tests/golden/extract_operator_goldens.py /root/reference
"""
import json
import os
import re
import sys


def parse(ref_root):
    base = os.path.join(ref_root, "tests", "integration", "operator")
    out = []
    for dirpath, _, files in sorted(os.walk(base)):
        if "expected" not in files or "opts" not in files:
            continue
        opts = {}
        with open(os.path.join(dirpath, "opts")) as f:
            for line in f:
                line = line.strip()
                if not line.startswith("-"):
                    continue
                parts = line.split()
                opts[parts[0][1:]] = parts[1] if len(parts) > 1 else True
        vals = {}
        test = None
        with open(os.path.join(dirpath, "expected")) as f:
            for line in f:
                m = re.search(r"Test \[(\w+)\]: Mesh (\d+) x (\d+) x (\d+)", line)
                if m:
                    test = m.group(1)
                    continue
                m = re.search(r"^\s+y\.y\s+=\s+([-+]?\S+)\s+\[mfo\]", line)
                if m and test:
                    v = m.group(1)
                    vals[test] = None if "nan" in v else float(v)
        if not vals:
            continue
        out.append({
            "source": os.path.relpath(os.path.join(dirpath, "expected"), ref_root),
            "mx": int(opts["mx"]), "my": int(opts["my"]), "mz": int(opts["mz"]),
            "eta1": float(opts.get("model_viscous_sinker_eta1", 1.0)),
            "a11_op": opts.get("a11_op", "default"),
            "yy": vals,
        })
    return out


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    data = parse(ref)
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "operator_goldens.json"), "w") as f:
        json.dump(data, f, indent=1)
    print(len(data), "golden files")
