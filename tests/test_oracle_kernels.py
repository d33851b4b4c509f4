"""Pin the oracle's Fortran-kernel restatement (CPU only).

* bit-for-bit against the reference's own Fortran compiled into oracle/_ref
  (when that library has been built here);
* against dense linear algebra / numpy.polynomial for what the mathematics fixes.
"""
import ctypes
import os

import numpy as np
import numpy.polynomial.chebyshev as npcheb
import pytest

from oracle import pypde_oracle as o

REF_LIB = os.path.join(os.path.dirname(o.__file__), "_ref", "libpypde_f90.so")


def _ref():
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref not built (reference tree absent)")
    lib = ctypes.CDLL(REF_LIB)
    dp, i = ctypes.POINTER(ctypes.c_double), ctypes.c_int
    lib.ref_diff_2d.argtypes = [dp, dp, i, i]
    lib.ref_tdma_2d.argtypes = [dp, dp, dp, dp, i, dp, i, i]
    lib.ref_fdma_2d.argtypes = [dp, dp, dp, dp, dp, i, i, i]
    lib.ref_init_fdma.argtypes = [dp, dp, dp, dp, dp, i]
    lib.ref_fdma_type2.argtypes = [dp, dp, dp, dp, i, i, i, i]
    lib.ref_twodma_2d.argtypes = [dp, dp, dp, i, i, i]
    return lib


def P(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


rng = np.random.default_rng(11)


def test_diff_matches_reference_fortran():
    lib = _ref()
    c = rng.standard_normal((37, 5))
    cf = np.asfortranarray(c)
    dc = np.zeros((37, 5), order="F")
    lib.ref_diff_2d(P(cf), P(dc), 37, 5)
    assert np.array_equal(o.diff_2d(c), dc)


def test_diff_matches_chebder():
    c = rng.standard_normal(30)
    d = o.diff_2d(c[:, None])[:, 0]
    ref = np.zeros(30)
    ref[:29] = npcheb.chebder(c)
    assert np.allclose(d, ref, rtol=1e-13, atol=1e-12 * np.abs(ref).max())
    # a plausible index slip (recurrence shifted by one) would fail this
    assert d[-1] == 0.0 and d[0] == pytest.approx(ref[0], rel=1e-13)


def test_tdma_matches_reference_and_dense():
    lib = _ref()
    n, m = 40, 3
    a, b, c = rng.random(n - 2), 2 + rng.random(n), rng.random(n - 2)
    d = rng.standard_normal((n, m))
    x = o.tdma_solve(a, b, c, d, 2)
    xr = np.zeros((n, m), order="F")
    lib.ref_tdma_2d(P(a.copy()), P(b.copy()), P(c.copy()), P(np.asfortranarray(d)), 2, P(xr), n, m)
    assert np.array_equal(x, xr)
    A = np.diag(b) + np.diag(a, -2) + np.diag(c, 2)
    assert np.allclose(x, np.linalg.solve(A, d), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("axis", [0, 1])
def test_fdma_matches_reference_and_dense(axis):
    lib = _ref()
    n = 23
    A = (np.diag(3 + rng.random(n)) + np.diag(rng.random(n - 2), -2) +
         np.diag(rng.random(n - 2), 2) + np.diag(rng.random(n - 4), 4))
    d, u1, u2, l = o.init_fdma(A)
    dr, u1r, u2r, lr = np.zeros(n), np.zeros(n - 2), np.zeros(n - 4), np.zeros(n - 2)
    lib.ref_init_fdma(P(np.asfortranarray(A)), P(dr), P(u1r), P(u2r), P(lr), n)
    assert np.array_equal(d, dr) and np.array_equal(u1, u1r) and np.array_equal(l, lr)
    # plan-time Python LU (plans.py:226-234) gives the same factors
    l2, d2, u12, u22 = np.diag(A, -2).copy(), np.diag(A).copy(), np.diag(A, 2).copy(), np.diag(A, 4).copy()
    o.fdma_lu(l2, d2, u12, u22)
    assert np.array_equal(d2, d) and np.array_equal(l2, l) and np.array_equal(u12, u1)
    b = rng.standard_normal((n, 7) if axis == 0 else (7, n))
    x = o.fdma_solve(l, d, u1, u2, b.copy(), axis)
    xr = np.asfortranarray(b.copy())
    lib.ref_fdma_2d(P(l), P(d), P(u1), P(u2), P(xr), axis, b.shape[0], b.shape[1])
    assert np.array_equal(x, xr)
    dense = np.linalg.solve(A, b if axis == 0 else b.T)
    assert np.allclose(x, dense if axis == 0 else dense.T, rtol=1e-12, atol=1e-12)


def test_type2_matches_reference_and_dense():
    lib = _ref()
    n, m = 18, 6
    A = np.diag(rng.random(n)) + np.diag(rng.random(n - 2), 2)
    A[0, 0] = 0.0
    C = (np.diag(2 + rng.random(n)) + np.diag(0.1 * rng.random(n - 2), -2) +
         np.diag(0.1 * rng.random(n - 2), 2) + np.diag(0.1 * rng.random(n - 4), 4))
    lam = np.concatenate([[1e-20], rng.random(m - 1) + 0.5])
    b = rng.standard_normal((n, m))
    x = o.fdma_type2(A, C, lam, b.copy(), 0, True)
    xr = np.asfortranarray(b.copy())
    lib.ref_fdma_type2(P(np.asfortranarray(A)), P(np.asfortranarray(C)), P(lam), P(xr), 0, 1, n, m)
    assert np.array_equal(x, xr)
    for j in range(1, m):
        assert np.allclose(x[:, j], np.linalg.solve(A + lam[j] * C, b[:, j]), rtol=1e-12, atol=1e-12)
    red = np.linalg.solve((A + lam[0] * C)[1:, 1:], b[1:, 0])
    assert x[0, 0] == 0.0 and np.allclose(x[1:, 0], red, rtol=1e-12, atol=1e-12)


def test_twodma_matches_reference_and_dense():
    lib = _ref()
    n = 15
    d, u = 1 + rng.random(n), rng.random(n - 2)
    b = rng.standard_normal((n, 4))
    x = o.twodma_solve(d, u, b.copy(), 0)
    xr = np.asfortranarray(b.copy())
    lib.ref_twodma_2d(P(d), P(u), P(xr), 0, n, 4)
    assert np.array_equal(x, xr)
    assert np.allclose(x, np.linalg.solve(np.diag(d) + np.diag(u, 2), b), rtol=1e-12)


def test_pdma_matches_dense():
    n = 20
    A = (np.diag(4 + rng.random(n)) + np.diag(rng.random(n - 2), -2) + np.diag(rng.random(n - 2), 2) +
         np.diag(rng.random(n - 4), -4) + np.diag(rng.random(n - 4), 4))
    tabs = o.pdma_factor(A)
    b = rng.standard_normal(n)
    x = o.pdma_solve(*tabs, b.copy())
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-12)
