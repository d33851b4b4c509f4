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


@pytest.mark.parametrize("variant", range(1, 43))
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


def _sym_operator(rng, nb, Lb, K, signs):
    """Rows in blocks of Lb, mirror-symmetric within each block:
    G[b*Lb + Lb-1-j, k] = signs[b] (-1)^k G[b*Lb + j, k]."""
    G = np.zeros((nb * Lb, K))
    Lh = (Lb + 1) // 2
    par = (-1.0) ** np.arange(K)
    for b in range(nb):
        top = rng.standard_normal((Lh, K))
        if Lb % 2:  # middle row: only the components the symmetry allows
            top[-1] *= (signs[b] * par == 1)
        G[b * Lb:b * Lb + Lh] = top
        for j in range(Lb - Lh):
            G[b * Lb + Lb - 1 - j] = signs[b] * par * top[j]
    return G


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5, 6, 7])
@pytest.mark.parametrize("Lb,K", [(8, 10), (7, 9), (96, 62)])
def test_gemm_output_fold(amd, variant, Lb, K):
    """fold 1 (rows) and 2 (columns): C = G X and C = Y G^T for mirror-symmetric
    G from half-size products over even / odd k."""
    import ctypes
    _, ops, torch = amd
    from pypde_amd._lib import lib
    rng = np.random.default_rng(Lb + K + variant)
    nb, n, batch = 2, 37, 3
    signs = [1.0, -1.0]
    G = _sym_operator(rng, nb, Lb, K, signs)
    Lh = (Lb + 1) // 2
    top = np.concatenate([G[b * Lb:b * Lb + Lh] for b in range(nb)])
    Ge, Go = ops.to_dev(np.ascontiguousarray(top[:, 0::2])), ops.to_dev(np.ascontiguousarray(top[:, 1::2]))
    ke, ko = (K + 1) // 2, K // 2
    X = rng.standard_normal((batch, K, n))
    dX = ops.to_dev(X)
    C = torch.empty((batch, nb * Lb, n), dtype=torch.float64, device=dX.device)
    d = ops._desc(0, 0, nb * Lh, n, ke, batch, 1.0, 0.0, Ge.data_ptr(), ke, 0, dX.data_ptr(), 2 * n,
                  dX.stride(0), None, 0, 0, C.data_ptr(), n, C.stride(0))
    d.fold, d.A1, d.B1, d.lda1, d.ldb1, d.k1 = 1, Go.data_ptr(), dX.data_ptr() + 8 * n, ko, 2 * n, ko
    d.fold_len, d.fold_sign, d.variant = Lb, 0b10, variant
    assert lib.pypde_dgemm_batched(ctypes.byref(d), ops.stream_ptr()) == 0
    assert rel(host(C), np.einsum("ik,bkn->bin", G, X)) < 1e-14
    # columns: C = Y G^T
    Y = rng.standard_normal((batch, n, K))
    dY = ops.to_dev(Y)
    C2 = torch.empty((batch, n, nb * Lb), dtype=torch.float64, device=dY.device)
    d = ops._desc(0, 1, n, nb * Lh, ke, batch, 1.0, 0.0, dY.data_ptr(), K, dY.stride(0),
                  Ge.data_ptr(), ke, 0, None, 0, 0, C2.data_ptr(), nb * Lb, C2.stride(0))
    d.fold, d.A1, d.B1, d.lda1, d.ldb1, d.k1 = 2, dY.data_ptr() + 8, Go.data_ptr(), K, ko, ko
    d.ska, d.fold_len, d.fold_sign, d.variant = 2, Lb, 0b10, variant
    assert lib.pypde_dgemm_batched(ctypes.byref(d), ops.stream_ptr()) == 0
    assert rel(host(C2), np.einsum("bnk,ik->bni", Y, G)) < 1e-14


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5])
def test_gemm_input_fold(amd, variant):
    """fold 3 / 4: X[k] + s_e X[Kf-1-k] folded into A (3) or B (4) with the sign
    alternating over batch entries; outputs interleaved with column stride 2."""
    import ctypes
    _, ops, torch = amd
    from pypde_amd._lib import lib
    rng = np.random.default_rng(40 + variant)
    Kf, m, n, batch = 24, 19, 13, 4
    Kh = Kf // 2
    # fold 3: C_e = sum_k (A[:, k] + s_e A[:, Kf-1-k]) B_e[k, :], entries e = 2q + p
    A = rng.standard_normal((batch // 2, m, Kf))
    B = rng.standard_normal((2, Kh, n))
    dA, dB = ops.to_dev(A), ops.to_dev(B)
    C = torch.zeros((batch // 2, m, 2 * n), dtype=torch.float64, device=dA.device)
    d = ops._desc(0, 0, m, n, Kh, batch, 1.0, 0.0, dA.data_ptr(), Kf, 0, dB.data_ptr(), n, 0,
                  None, 0, 0, C.data_ptr(), 2 * n, 0)
    d.iA = ops._bi(m * Kf, d1=2)
    d.iB = ops._bi(Kh * n, m1=2)
    d.iC = ops._bi(m * 2 * n, d1=2, s2=1, m2=2)
    d.csn = 2
    d.fold, d.fold_len, d.mirror_sign, d.mirror_div, d.variant = 3, Kf, -1.0, 1, variant
    assert lib.pypde_dgemm_batched(ctypes.byref(d), ops.stream_ptr()) == 0
    ref = np.zeros((batch // 2, m, 2 * n))
    for e in range(batch):
        qq, pp = divmod(e, 2)
        s = -1.0 if pp == 0 else 1.0
        Af = A[qq][:, :Kh] + s * A[qq][:, ::-1][:, :Kh]
        ref[qq][:, pp::2] = Af @ B[pp]
    assert rel(host(C), ref) < 1e-14
    # fold 4: C_e = A_e (B[k] + s_e B[Kf-1-k]) with B transposed ([n][Kf]); rows interleaved
    At = rng.standard_normal((2, m, Kh))
    Bt = rng.standard_normal((batch // 2, n, Kf))
    dA, dB = ops.to_dev(At), ops.to_dev(Bt)
    C = torch.zeros((batch // 2, 2 * m, n), dtype=torch.float64, device=dA.device)
    d = ops._desc(0, 1, m, n, Kh, batch, 1.0, 0.0, dA.data_ptr(), Kh, 0, dB.data_ptr(), Kf, 0,
                  None, 0, 0, C.data_ptr(), 2 * n, 0)
    d.iA = ops._bi(m * Kh, m1=2)
    d.iB = ops._bi(n * Kf, d1=2)
    d.iC = ops._bi(2 * m * n, d1=2, s2=n, m2=2)
    d.fold, d.fold_len, d.mirror_sign, d.mirror_div, d.variant = 4, Kf, 1.0, 1, variant
    assert lib.pypde_dgemm_batched(ctypes.byref(d), ops.stream_ptr()) == 0
    ref = np.zeros((batch // 2, 2 * m, n))
    for e in range(batch):
        qq, pp = divmod(e, 2)
        s = 1.0 if pp == 0 else -1.0
        Bf = Bt[qq][:, :Kh] + s * Bt[qq][:, ::-1][:, :Kh]
        ref[qq][pp::2] = At[pp] @ Bf.T
    assert rel(host(C), ref) < 1e-14


@pytest.mark.parametrize("N", [23, 24, 64])
@pytest.mark.parametrize("axis", [0, 1])
def test_mirror_operator_transforms(amd, N, axis):
    """Folded transforms (ops.MirrorOperator) along both axes equal the dense
    products: Chebyshev backward to the dealiased grid (output fold), forward
    from it (input fold, even sizes), and a non-symmetric matrix (plain GEMM)."""
    pypde_amd, ops, torch = amd
    from pypde_amd.bases import cheb_backward_matrix, cheb_forward_matrix
    rng = np.random.default_rng(N + 10 * axis)
    L = int(1.5 * N)
    Bk = cheb_backward_matrix(L, N)      # [L, N]
    Fw = cheb_forward_matrix(L, N)       # [N, L]
    R = rng.standard_normal((N + 1, L))
    for M, want in ((Bk, 1), (Fw, 2 if (L % 2 == 0 and N % 2 == 0) else 0), (R, 0),
                    (cheb_backward_matrix(N, N), 2 if N % 2 == 0 else 1)):
        op = ops.MirrorOperator(M)
        assert op.mode == want
        n_in = M.shape[1]
        x = rng.standard_normal((3, n_in, 7) if axis == 0 else (3, 7, n_in))
        y = host(op.apply(ops.to_dev(x), axis))
        ref = np.einsum("ik,bkn->bin", M, x) if axis == 0 else np.einsum("bnk,ik->bni", x, M)
        assert rel(y, ref) < 1e-13, (M.shape, axis)


PLAIN_VARIANTS = tuple(range(1, 43))


def test_gemm_prefetch_depth_is_bit_identical(amd):
    """Variants with two slabs of loads in flight run the same slabs in the
    same order as their one-slab forms, so the results agree bit for bit:
    12..15 against 8..11 (split-K), 7, 24, 25, 26, 27 against 5, 6, 1, 3, 2; the
    double-buffered forms 4 and 28 against 3 and 2.  The 32x32 tile over fewer
    waves (29..32) runs the same MFMA chain per output, so it matches 2 too, and
    so do 33..36 (32- and 64-deep K slabs: the same chain, more zero padding);
    37 and 38 match 1 and 5 the same way; 39..42 (another LDS row pitch)
    match 2, 27, 33 and 34."""
    import ctypes
    _, ops, torch = amd
    from pypde_amd._lib import lib
    rng = np.random.default_rng(77)
    for (m, n, k, batch) in [(37, 53, 29, 3), (254, 254, 894, 2), (64, 64, 700, 1), (5, 3, 1, 2),
                             (100, 40, 65, 1)]:
        for ta in (0, 1):
            for tb in (0, 1):
                A = rng.standard_normal((batch, k, m) if ta else (batch, m, k))
                B = rng.standard_normal((batch, k, n) if not tb else (batch, n, k))
                dA, dB = ops.to_dev(A), ops.to_dev(B)
                outs = {}
                for v in (8, 9, 10, 11, 12, 13, 14, 15, 5, 7, 6, 24, 1, 25, 3, 26, 4, 2, 27, 28, 29, 30,
                          31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42):
                    C = torch.empty((batch, m, n), dtype=torch.float64, device=dA.device)
                    d = ops._desc(ta, tb, m, n, k, batch, 1.0, 0.0, dA.data_ptr(), dA.stride(1),
                                  dA.stride(0), dB.data_ptr(), dB.stride(1), dB.stride(0), None, 0,
                                  0, C.data_ptr(), n, m * n)
                    d.variant = v
                    assert lib.pypde_dgemm_batched(ctypes.byref(d), ops.stream_ptr()) == 0
                    outs[v] = host(C)
                for v, w in ((8, 12), (9, 13), (10, 14), (11, 15), (5, 7), (6, 24), (1, 25),
                             (3, 26), (3, 4), (2, 27), (2, 28), (2, 29), (2, 30), (2, 31),
                             (2, 32), (2, 33), (2, 34), (2, 35), (2, 36), (1, 37),
                             (5, 38), (2, 39), (27, 40), (33, 41), (34, 42)):
                    assert np.array_equal(outs[v], outs[w]), (m, n, k, ta, tb, v, w)


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_grouped_matches_members(amd, ta, tb):
    """pypde_dgemm_grouped: every member equals its own batched launch bit for
    bit, for each plain variant; members differ in shape, batch and epilogue;
    one member is empty."""
    import ctypes
    pypde_amd, ops, torch = amd
    from pypde_amd._lib import GemmDesc, lib
    rng = np.random.default_rng(40 + 2 * ta + tb)
    shapes = [(37, 53, 29, 3), (130, 70, 91, 1), (0, 16, 8, 2), (64, 200, 254, 2)]
    ops_, outs_single, outs_group, descs = [], [], [], []
    for (m, n, k, b) in shapes:
        A = ops.to_dev(rng.standard_normal((b, k, m) if ta else (b, m, k)))
        B = ops.to_dev(rng.standard_normal((b, n, k) if tb else (b, k, n)))
        D = ops.to_dev(rng.standard_normal((b, m, n)))
        av = ops.to_dev(rng.uniform(0.5, 1.5, b))
        ops_.append((A, B, D, av))
        outs_single.append(torch.zeros((b, m, n), dtype=torch.float64, device="cuda"))
        outs_group.append(torch.zeros((b, m, n), dtype=torch.float64, device="cuda"))
    s = ops.stream_ptr()
    for v in PLAIN_VARIANTS:
        arr = (GemmDesc * len(shapes))()
        for i, ((m, n, k, b), (A, B, D, av)) in enumerate(zip(shapes, ops_)):
            def mk(C):
                d = ops._desc(ta, tb, m, n, k, b, 0.75, -1.5, A.data_ptr(), m if ta else k,
                              A[0].numel(), B.data_ptr(), k if tb else n, B[0].numel(),
                              D.data_ptr(), n, m * n, C.data_ptr(), n, m * n,
                              alpha_vec=av.data_ptr())
                d.zero_first = 1 if i == 1 else 0
                d.variant = v
                return d
            d1 = mk(outs_single[i])
            ops.check(lib.pypde_dgemm_batched(ctypes.byref(d1), s), "single")
            arr[i] = mk(outs_group[i])
        ops.check(lib.pypde_dgemm_grouped(ctypes.byref(arr), len(shapes), s), "grouped")
        torch.cuda.synchronize()
        for i in range(len(shapes)):
            assert torch.equal(outs_single[i], outs_group[i]), (v, i)
            outs_group[i].zero_()
    # the results themselves
    (m, n, k, b), (A, B, D, av) = shapes[0], ops_[0]
    An, Bn, Dn, avn = (host(t) for t in (A, B, D, av))
    for e in range(b):
        ref = 0.75 * avn[e] * ((An[e].T if ta else An[e]) @ (Bn[e].T if tb else Bn[e])) - 1.5 * Dn[e]
        assert rel(host(outs_single[0][e]), ref) < 1e-14


def test_gemm_grouped_rejects_mixed_members(amd):
    import ctypes
    pypde_amd, ops, torch = amd
    from pypde_amd._lib import GemmDesc, PypdeError, lib
    x = torch.zeros((8, 8), dtype=torch.float64, device="cuda")
    p = x.data_ptr()

    def mk(ta=0, tb=0, v=2, fold=0):
        d = ops._desc(ta, tb, 8, 8, 8, 1, 1.0, 0.0, p, 8, 0, p, 8, 0, None, 0, 0, p, 8, 0)
        d.variant, d.fold = v, fold
        return d
    s = ops.stream_ptr()
    for pair in ((mk(), mk(tb=1)), (mk(), mk(v=9)), (mk(), mk(fold=1))):
        arr = (GemmDesc * 2)(*pair)
        with pytest.raises(PypdeError):
            ops.check(lib.pypde_dgemm_grouped(ctypes.byref(arr), 2, s), "grouped")
    arr = (GemmDesc * 5)(*[mk() for _ in range(5)])
    for count in (0, 5):
        with pytest.raises(PypdeError):
            ops.check(lib.pypde_dgemm_grouped(ctypes.byref(arr), count, s), "grouped")


@pytest.mark.parametrize("axis", [0, 1])
def test_banded_every_row_hint_bit_exact(amd, axis):
    """Every row hint of the banded and stencil applies (1..8 outputs per
    thread / rows per block, 18/20/24 the register-window forms, 33..40 the
    coefficient-resident form along rows, with row counts that leave partial
    groups of four and partial blocks) gives results
    bit-identical to the oracle: B (offsets 0, 2, 4), B S, a band with negative
    offsets and a gap, and the ChebDirichlet / ChebNeumann stencils."""
    pypde_amd, ops, torch = amd
    from oracle import pypde_oracle as o
    rng = np.random.default_rng(50 + axis)
    N, other = 41, 37
    hints = (1, 2, 4, 8, 18, 20, 24) + (tuple(range(33, 41)) if axis == 1 else ())
    ob = o.Basis(N, "CD")
    B, BS, _ = o.hholtz_matrices(ob, 0.02, 0.5)
    G = np.zeros((N - 3, N))
    for i in range(N - 3):
        for off in (-4, -1, 0, 3):
            if 0 <= i + off < N:
                G[i, i + off] = rng.standard_normal()
    for Mat in (B, BS, G):
        nin = Mat.shape[1]
        x = rng.standard_normal((nin, other) if axis == 0 else (other, nin))
        ref = o.MatAxis(Mat, axis)(x)
        p = pypde_amd.PlanRHS(Mat, ndim=2, axis=axis)
        assert p.kind == "banded"
        for hint in hints:
            y = host(ops.banded_apply(p.offsets, p.vals, ops.to_dev(x), axis, p.M_out, lpb=hint))
            assert np.array_equal(y, ref), (Mat.shape, hint)
        # a batch of three arrays
        x3 = rng.standard_normal((3,) + x.shape)
        ref3 = np.stack([o.MatAxis(Mat, axis)(a) for a in x3])
        for hint in hints:
            y = host(ops.banded_apply(p.offsets, p.vals, ops.to_dev(x3), axis, p.M_out,
                                      lpb=hint))
            assert np.array_equal(y, ref3), (Mat.shape, hint, "batch")
    for key in ("CD", "CN"):
        b, obk = pypde_amd.Base(N, key), o.Basis(N, key)
        v = rng.standard_normal((N - 2, other) if axis == 0 else (other, N - 2))
        ref = o._along(obk.to_chebyshev, v, axis)
        for hint in hints:
            u = host(ops.stencil_apply(b.stencil_vector(), ops.to_dev(v), axis, lpb=hint))
            assert np.array_equal(u, ref), (key, hint)
    if axis == 0:
        from pypde_amd._lib import PypdeError
        with pytest.raises(PypdeError):
            ops.banded_apply(p.offsets, p.vals, ops.to_dev(x), 0, p.M_out, lpb=33)


def test_split_fold_variants_reject_input_folds(amd):
    """Fold variants 6 and 7 (each product over its own waves) serve the output
    folds only; an input-fold descriptor asking for them is rejected before
    any launch."""
    import ctypes
    _, ops, torch = amd
    from pypde_amd._lib import lib
    x = torch.zeros((64, 64), dtype=torch.float64, device="cuda")
    p = x.data_ptr()
    for fold, v in ((3, 6), (4, 7)):
        d = ops._desc(0, 0, 8, 8, 4, 1, 1.0, 0.0, p, 8, 0, p, 8, 0, None, 0, 0,
                      p + 8 * 32 * 64, 8, 0)
        d.fold, d.fold_len, d.mirror_sign, d.mirror_div, d.variant = fold, 8, 1.0, 1, v
        assert lib.pypde_dgemm_batched(ctypes.byref(d), ops.stream_ptr()) != 0


@pytest.mark.parametrize("N", [9, 24, 64, 65, 128, 256, 512])
@pytest.mark.parametrize("axis", [0, 1])
def test_lds_dct_every_form_matches_oracle(amd, N, axis):
    """The LDS-staged cosine-sum transform (pypde_cheb_dct) in each of its
    kernel forms, and the MFMA form, against the oracle's scipy DCT-I
    (chebyshev.py:67-93), backward and forward, on ragged batches along both
    axes (odd and even N: with and without a middle output pair)."""
    pypde_amd, ops, torch = amd
    from oracle import pypde_oracle as o
    rng = np.random.default_rng(N + 31 * axis)
    b = pypde_amd.Base(N, "CH")
    ob = o.Basis(N, "CH")
    other = 37
    shape = (3, N, other) if axis == 0 else (3, other, N)
    x = rng.standard_normal(shape)
    for kind in ("bwd", "fwd"):
        op = b._transform(kind)
        fn = ob._cheb_backward if kind == "bwd" else ob._cheb_forward
        ref = np.stack([o._along(fn, x[e], axis) for e in range(3)])
        for form in op.FORMS:
            y = host(op.apply(ops.to_dev(x), axis, form=form))
            e = rel(y, ref)
            assert e < 1e-13, (kind, form, e)
