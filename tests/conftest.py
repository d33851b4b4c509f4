import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

GOLDEN = os.path.join(REPO, "tests", "golden", "rbc2d_reference.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (runs the pypde_amd kernels)")


@pytest.fixture(scope="session")
def golden():
    return np.load(GOLDEN, allow_pickle=False)


def rel(a, b):
    """Norm-wise relative difference ||a-b|| / ||b||."""
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def oracle_runs(make, nsteps):
    """(ref, dense): the oracle solver made by make() after nsteps updates, and
    the same run with the oracle's dense-product transforms (same arithmetic,
    different rounding; its distance from ref is the rounding floor)."""
    from oracle import pypde_oracle as o
    ref = make()
    for _ in range(nsteps):
        ref.update()
    with o.dense_transforms():
        dense = make()
        for _ in range(nsteps):
            dense.update()
    return ref, dense


def floor_tol(ref, dense, field, base):
    """Tolerance for a field: base, or twice the rounding floor where that is
    larger."""
    a, b = getattr(dense, field).vhat, getattr(ref, field).vhat
    return max(base, 2.0 * rel(a, b))
