import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (REPO, os.path.join(REPO, "ptatin-gene_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libptq2 on cuda:0)")


@pytest.fixture(scope="session")
def goldens():
    import json

    with open(os.path.join(REPO, "tests", "golden", "operator_goldens.json")) as f:
        return json.load(f)
