import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "phantom-grape_amd")
for p in (PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP libraries)")


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def rel_err(a, b, axis=-1, floor=0.0):
    """Per-row relative error |a-b| / max(|b|, floor) (vectors along `axis`)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.ndim == 1:
        return np.abs(a - b) / np.maximum(np.abs(b), floor)
    return np.linalg.norm(a - b, axis=axis) / np.maximum(np.linalg.norm(b, axis=axis), floor)


@pytest.fixture(scope="session")
def ref_tree():
    """Path of the reference sources if they are in this container."""
    p = os.environ.get("PGAMD_REFERENCE", "/root/reference")
    return p if os.path.isdir(p) else None
