import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libsfac.so)")
    config.addinivalue_line("markers", "slow: full-size configurations")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(GOLDEN, "test_log_ref.json")) as f:
        return json.load(f)


def case_T(golden, q=None):
    """The reference's 4-atom triclinic case (test.py:66-86)."""
    c = golden["case"]
    L = np.array(c["L"])
    crds = np.array(c["crds_frac"])
    from oracle.sfac_oracle import LofS
    atoms = crds @ LofS(L)
    q = np.array(c["q"]) if q is None else q
    return L, np.array(c["K"]), c["order"], crds, atoms, q


def rel_rms(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) /
                 max(np.max(np.abs(np.asarray(b))), 1e-300))
