"""pytest configuration: the `gpu` marker and shared fixtures.

GPU tests (marked gpu) call the HIP library through its C ABI and compare
against the oracle (oracle/, the checker) and the golden fixtures written
from the compiled reference (tests/golden/).
"""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X)")


@pytest.fixture(scope="session")
def golden():
    data = np.load(os.path.join(GOLDEN, "golden.npz"))
    with open(os.path.join(GOLDEN, "golden_cases.json")) as f:
        cases = json.load(f)
    return data, cases


def relerr(y, yref):
    """normwise relative error ||y - yref||_2 / ||yref||_2"""
    y = np.asarray(y, float).ravel()
    yref = np.asarray(yref, float).ravel()
    return float(np.linalg.norm(y - yref) / np.linalg.norm(yref))


def maxrel(y, yref):
    """max-abs error over the max-abs of the reference"""
    y = np.asarray(y, float).ravel()
    yref = np.asarray(yref, float).ravel()
    return float(np.abs(y - yref).max() / np.abs(yref).max())
