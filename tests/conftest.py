import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libpps.so")


def gpu_available():
    """A ROCm device node is present (no HIP call, no device initialisation)."""
    import glob
    return os.path.exists("/dev/kfd") and bool(glob.glob("/dev/dri/renderD*"))


@pytest.fixture(scope="session")
def golden():
    import json
    import numpy as np
    here = os.path.join(ROOT, "tests", "golden")
    cases = json.load(open(os.path.join(here, "ref_cases.json")))["cases"]
    vecs = np.load(os.path.join(here, "ref_vectors.npz"))
    return {c["name"]: c for c in cases}, vecs
