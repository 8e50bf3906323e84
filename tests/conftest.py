import os
import sys

import numpy as np
import pytest

try:
    # torch first: the HIP runtime it initialises must not find the device
    # already set up by libxs_hip, or torch reports no GPU (the GPU tests that
    # hand device buffers to the library use torch allocations)
    import torch  # noqa: F401
except ImportError:
    torch = None

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "softx-d-16-00083_amd")
for p in (PKG, os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def golden(name):
    return np.load(os.path.join(GOLDEN, f"ref_{name}.npz"), allow_pickle=False)


@pytest.fixture(scope="session")
def ctx():
    import xansons_hip as xh
    c = xh.Context(0)
    yield c
    c.close()


def rel_err(a, ref, floor_frac=1e-3):
    """Max pointwise relative error over points with |ref| >= floor_frac*max|ref|,
    and the L2-relative error."""
    a, ref = np.asarray(a, float), np.asarray(ref, float)
    m = np.abs(ref) >= floor_frac * np.abs(ref).max()
    pw = np.max(np.abs(a[m] - ref[m]) / np.abs(ref[m])) if m.any() else 0.0
    l2 = np.linalg.norm(a - ref) / max(np.linalg.norm(ref), 1e-300)
    return float(pw), float(l2)
