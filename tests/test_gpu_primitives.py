"""GPU parity of the C-ABI primitives against the oracle (each call goes
through libpypde_hip.so).  Banded line kernels run the reference's arithmetic
statement by statement and must be bit-identical to the oracle; GEMM-based
transforms reorder sums and are held to 1e-13 norm-wise relative error."""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amd():
    import torch
    import pypde_amd
    from pypde_amd import ops
    assert torch.cuda.is_available()
    return pypde_amd, ops, torch


def host(t):
    return t.detach().cpu().numpy()


def test_mfma_gemm_layout(amd):
    _, ops, torch = amd
    rng = np.random.default_rng(0)
    for (m, n, k) in [(16, 16, 4), (37, 53, 29), (130, 70, 91), (256, 256, 256)]:
        for ta in (False, True):
            for tb in (False, True):
                A = rng.standard_normal((k, m) if ta else (m, k))
                B = rng.standard_normal((n, k) if tb else (k, n))
                D = rng.standard_normal((m, n))
                C = ops.gemm(ops.to_dev(A), ops.to_dev(B), ta, tb, alpha=0.5, beta=-2.0,
                             D=ops.to_dev(D))
                ref = 0.5 * ((A.T if ta else A) @ (B.T if tb else B)) - 2.0 * D
                assert rel(host(C), ref) < 1e-14, (m, n, k, ta, tb)


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 8, 9, 10, 11, 16, 17, 18, 19, 20, 21, 22, 23])
def test_gemm_every_variant(amd, variant):
    """Each kernel variant of pypde_dgemm_batched on ragged shapes (m, n not
    multiples of the tile, k not a multiple of the split), all transposes,
    batch offsets with two-term indexing, per-entry alpha and beta * D."""
    import ctypes
    _, ops, torch = amd
    from pypde_amd._lib import lib
    rng = np.random.default_rng(variant)
    for (m, n, k, batch) in [(16, 16, 4, 1), (37, 53, 29, 3), (130, 70, 91, 2), (254, 254, 894, 3),
                             (5, 3, 1, 2), (64, 64, 700, 1)]:
        for ta in (0, 1):
            for tb in (0, 1):
                A = rng.standard_normal((batch, k, m) if ta else (batch, m, k))
                B = rng.standard_normal((k, n) if not tb else (n, k))  # broadcast over batch
                D = rng.standard_normal((batch, m, n))
                av = rng.standard_normal(batch)
                dA, dB, dD, dav = (ops.to_dev(t) for t in (A, B, D, av))
                C = torch.empty((batch, m, n), dtype=torch.float64, device=dA.device)
                d = ops._desc(ta, tb, m, n, k, batch, 0.5, -2.0, dA.data_ptr(), dA.stride(1),
                              dA.stride(0), dB.data_ptr(), dB.stride(0), 0, dD.data_ptr(),
                              dD.stride(1), dD.stride(0), C.data_ptr(), C.stride(1), C.stride(0),
                              dav.data_ptr())
                d.variant = variant
                assert lib.pypde_dgemm_batched(ctypes.byref(d), ops.stream_ptr()) == 0
                opA = np.swapaxes(A, 1, 2) if ta else A
                opB = B.T if tb else B
                ref = 0.5 * av[:, None, None] * (opA @ opB) - 2.0 * D
                assert rel(host(C), ref) < 1e-14, (variant, m, n, k, batch, ta, tb)


def test_gemm_batched_strided(amd):
    _, ops, torch = amd
    rng = np.random.default_rng(1)
    A = rng.standard_normal((40, 33))
    X = rng.standard_normal((5, 33, 21))
    out = ops.matmul_axis(ops.to_dev(A), ops.to_dev(X), axis=0)
    assert rel(host(out), np.einsum("ij,bjk->bik", A, X)) < 1e-14
    Y = rng.standard_normal((5, 21, 33))
    out = ops.matmul_axis(ops.to_dev(A), ops.to_dev(Y), axis=1)
    assert rel(host(out), np.einsum("bkj,ij->bki", Y, A)) < 1e-14


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("key", ["CD", "CN"])
def test_stencil_maps_bit_exact(amd, axis, key):
    pypde_amd, ops, torch = amd
    from oracle import pypde_oracle as o
    N, other = 34, 19
    rng = np.random.default_rng(2)
    b = pypde_amd.Base(N, key)
    ob = o.Basis(N, key)
    v = rng.standard_normal((N - 2, other) if axis == 0 else (other, N - 2))
    u = host(b.to_chebyshev(ops.to_dev(v), axis=axis))
    ref = o._along(ob.to_chebyshev, v, axis)
    assert np.array_equal(u, ref)
    uu = rng.standard_normal((N, other) if axis == 0 else (other, N))
    ref = o._along(ob.from_chebyshev, uu, axis)
    vv = host(b.from_chebyshev(ops.to_dev(uu), axis=axis, exact=True))
    assert np.array_equal(vv, ref)
    # register-resident kernel (reciprocal pivots): rounding-level agreement
    vv = host(b.from_chebyshev(ops.to_dev(uu), axis=axis))
    assert rel(vv, ref) < 1e-14


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("key", ["CH", "CD", "CN"])
def test_derivative_bit_exact(amd, axis, key):
    pypde_amd, ops, torch = amd
    from oracle import pypde_oracle as o
    N, other = 27, 11
    rng = np.random.default_rng(3)
    b = pypde_amd.Base(N, key)
    ob = o.Basis(N, key)
    M = N if key == "CH" else N - 2
    v = rng.standard_normal((M, other) if axis == 0 else (other, M))
    for k in (1, 2):
        d = host(b.derivative(ops.to_dev(v), k, axis=axis, scale_pow=0.5 ** k))
        ref = o._along(lambda a: ob.derivative(a, k), v, axis) / 0.5 ** k
        assert np.array_equal(d, ref), k


@pytest.mark.parametrize("N", [16, 64, 256])
@pytest.mark.parametrize("axis", [0, 1])
def test_fdma_bit_exact_and_transpose(amd, N, axis):
    pypde_amd, ops, torch = amd
    from oracle import pypde_oracle as o
    rng = np.random.default_rng(4)
    f = o.Field([o.Basis(N, "CD"), o.Basis(N, "CN")])
    plan = o.solverplan_hholtz2d_adi(f.xs, 0.01, (0.5, 0.5))
    A = plan.mats["Ax"] if axis == 0 else plan.mats["Ay"]
    p = pypde_amd.Plan_fdma(A, ndim=2, axis=axis)
    b = rng.standard_normal((N - 2, N - 2))
    ref = o.FdmaAxis(A, axis)(b)
    x = host(p.solve(ops.to_dev(b), exact=True))
    assert np.array_equal(x, ref)
    x = host(p.solve(ops.to_dev(b)))
    assert rel(x, ref) < 1e-14
    # transposed solve: (A^-T) against a dense solve
    xt = host(ops.fdma_solve(p.tabs, ops.to_dev(b), axis, transpose=True))
    dense = np.linalg.solve(A.T, b if axis == 0 else b.T)
    assert rel(xt, dense if axis == 0 else dense.T) < 1e-11


def test_banded_rhs_bit_exact(amd):
    pypde_amd, ops, torch = amd
    from oracle import pypde_oracle as o
    rng = np.random.default_rng(5)
    N = 40
    ob = o.Basis(N, "CD")
    B, BS, _ = o.hholtz_matrices(ob, 0.02, 0.5)
    for axis in (0, 1):
        for Mat, nin in ((B, N), (BS, N - 2)):
            x = rng.standard_normal((nin, 13) if axis == 0 else (13, nin))
            y = host(pypde_amd.PlanRHS(Mat, ndim=2, axis=axis).solve(ops.to_dev(x)))
            ref = o.MatAxis(Mat, axis)(x)
            assert np.array_equal(y, ref)


def test_poisson_type2_bit_exact(amd):
    pypde_amd, ops, torch = amd
    from oracle import pypde_oracle as o
    rng = np.random.default_rng(6)
    for shape in ((18, 14), (66, 66)):
        f = o.Field([o.Basis(shape[0], "CN"), o.Basis(shape[1], "CN")])
        m = o.poisson_matrices(f.xs, singular=True, scale=(0.5, 0.5))
        p = pypde_amd.Plan_Poisson(m["Ax"], m["wy"], m["Cx"], ndim=2, axis=0, singular=True)
        b = rng.standard_normal((shape[0] - 2, shape[1] - 2))
        ref = o.fdma_type2(m["Ax"], m["Cx"], m["wy"], b.copy(), 0, True)
        x = host(p.solve(ops.to_dev(b), exact=True))
        assert np.array_equal(x, ref)
        x = host(p.solve(ops.to_dev(b)))
        assert rel(x, ref) < 1e-13 and x[0, 0] == 0.0
        xt = host(ops.poisson_solve(p.tab, p.sing, ops.to_dev(b), 0, transpose=True))
        for j in range(1, b.shape[1]):
            Aj = m["Ax"] + m["wy"][j] * m["Cx"]
            assert rel(xt[:, j], np.linalg.solve(Aj.T, b[:, j])) < 1e-10


def test_twodma_and_pdma(amd):
    pypde_amd, ops, torch = amd
    from oracle import pypde_oracle as o
    rng = np.random.default_rng(7)
    n = 30
    A = np.diag(1 + rng.random(n)) + np.diag(rng.random(n - 2), 2)
    b = rng.standard_normal((n, 5))
    pl = pypde_amd.Plan_twodma(A, ndim=2, axis=0)
    x = host(pl.solve(ops.to_dev(b), exact=True))
    assert np.array_equal(x, o.TwodmaAxis(A, 0)(b))
    x = host(pl.solve(ops.to_dev(b)))
    assert rel(x, o.TwodmaAxis(A, 0)(b)) < 1e-14
    P5 = (np.diag(4 + rng.random(n)) + np.diag(rng.random(n - 2), -2) + np.diag(rng.random(n - 2), 2) +
          np.diag(rng.random(n - 4), -4) + np.diag(rng.random(n - 4), 4))
    pd = pypde_amd.PDMA(P5)
    pd.init()
    x = host(pd.solve(ops.to_dev(b)))
    ref = o.pdma_solve(*o.pdma_factor(P5), b.copy())
    assert np.array_equal(x, ref)
    assert rel(x, np.linalg.solve(P5, b)) < 1e-12


@pytest.mark.parametrize("N", [24, 64, 256])
@pytest.mark.parametrize("key", ["CH", "CD", "CN"])
def test_transforms_match_oracle(amd, N, key):
    pypde_amd, ops, torch = amd
    from oracle import pypde_oracle as o
    rng = np.random.default_rng(8)
    b = pypde_amd.Base(N, key)
    ob = o.Basis(N, key)
    f = rng.standard_normal((N, 9))
    c = host(b.forward_fft(ops.to_dev(f)))
    cref = ob.forward_fft(f)
    assert rel(c, cref) < 1e-13
    u = host(b.backward_fft(ops.to_dev(cref)))
    assert rel(u, ob.backward_fft(cref)) < 1e-13


@pytest.mark.parametrize("N", [8, 64, 200, 256, 512])
@pytest.mark.parametrize("axis", [0, 1])
def test_recurrence_chunking_lengths(amd, N, axis):
    """Line lengths around the 64-position chunk boundaries and several
    batch entries / line counts that are not multiples of 64."""
    pypde_amd, ops, torch = amd
    from oracle import pypde_oracle as o
    rng = np.random.default_rng(N + axis)
    f = o.Field([o.Basis(N, "CD"), o.Basis(N, "CD")])
    plan = o.solverplan_hholtz2d_adi(f.xs, 0.05, (0.5, 0.5))
    A = plan.mats["Ax"]
    p = pypde_amd.Plan_fdma(A, ndim=2, axis=axis)
    M = N - 2
    other = 37
    b = rng.standard_normal((3, M, other) if axis == 0 else (3, other, M))
    x = host(p.solve(ops.to_dev(b)))
    for e in range(3):
        ref = o.FdmaAxis(A, axis)(b[e])
        assert rel(x[e], ref) < 1e-13, e
