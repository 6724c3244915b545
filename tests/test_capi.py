"""Compile a C consumer of the C ABI with gcc and run it (no GPU needed)."""
import os
import subprocess

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "ptatin-gene_amd")


def test_c_consumer(tmp_path):
    exe = str(tmp_path / "test_capi")
    subprocess.check_call(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(REPO, "include"),
                           os.path.join(REPO, "tests", "c", "test_capi.c"), "-L", PKG, "-lptq2",
                           "-Wl,-rpath," + PKG, "-o", exe])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr + out.stdout
    assert "capi ok" in out.stdout
