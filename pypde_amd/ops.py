"""Along-axis operators on device tensors, each one call into libpypde_hip.so.

Tensors are float64 CUDA (HIP) tensors of shape [n] , [n0, n1] or
[batch, n0, n1] with contiguous rows.  `axis` counts within the trailing two
dimensions (0 = columns, 1 = rows), like the reference's numpy axis argument.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import Lines, check, lib, ptr, stream_ptr

DTYPE = torch.float64


def device():
    _lib.require_gpu()
    return torch.device("cuda", torch.cuda.current_device())


def to_dev(a):
    """numpy / list / tensor -> contiguous float64 device tensor."""
    if isinstance(a, torch.Tensor):
        t = a.to(device=device(), dtype=DTYPE)
    else:
        arr = np.asarray(a, dtype=np.float64)
        if not arr.flags.writeable:  # read-only arrays (np.load) cannot back a tensor
            arr = arr.copy()
        t = torch.as_tensor(arr, device=device())
    return t.contiguous()


def _as3(t):
    """View a 1-D/2-D/3-D tensor as [batch, n0, n1] (1-D -> [1, n, 1])."""
    if t.dim() == 1:
        return t.reshape(1, t.shape[0], 1)
    if t.dim() == 2:
        return t.unsqueeze(0)
    if t.dim() == 3:
        return t
    raise NotImplementedError("ndim must be less than 4")


def _geom(src3, dst3, axis, len_in, len_out, lpb=0):
    if src3.stride(2) != 1 or dst3.stride(2) != 1:
        raise ValueError("rows must be contiguous")
    g = Lines()
    g.src = src3.data_ptr()
    g.dst = dst3.data_ptr()
    g.src_ld = src3.stride(1)
    g.dst_ld = dst3.stride(1)
    g.src_bs = src3.stride(0) if src3.shape[0] > 1 else 0
    g.dst_bs = dst3.stride(0) if dst3.shape[0] > 1 else 0
    g.len_in = len_in
    g.len_out = len_out
    g.nlines = src3.shape[2] if axis == 0 else src3.shape[1]
    g.batch = src3.shape[0]
    g.axis = axis
    g.lines_per_block = lpb
    return g


def _out_shape(x, axis, n_out):
    shape = list(x.shape)
    if x.dim() == 1:
        shape[0] = n_out
    else:
        shape[x.dim() - 2 + axis] = n_out
    return shape


def _len(x, axis):
    return x.shape[0] if x.dim() == 1 else x.shape[x.dim() - 2 + axis]


def _check_axis(x, axis):
    if x.dim() == 1 and axis != 0:
        raise ValueError("1-D input only has axis 0")
    if axis not in (0, 1):
        raise ValueError("axis must be 0 or 1")


# ---------------------------------------------------------------------------
def _bi(s1, d1=1, m1=0, s2=0, d2=1, m2=0):
    from ._lib import BIdx
    b = BIdx()
    b.s1, b.d1, b.m1, b.s2, b.d2, b.m2 = s1, d1, m1, s2, d2, m2
    return b


# ---------------------------------------------------------------------------
# Launch recording.  While a recorder is active, the GEMM launches made here
# (single descriptors and groups, each with the kernel variant the tuner chose)
# and the library calls made through call() are appended to it as well, so
# that a sequence of launches can be replayed without its Python logic or
# recorded into a step plan of the C-ABI.
_recorders = []


class Recorded:
    """One recorded launch: kind "gemm" (one descriptor), "group" (descriptors
    of one grouped launch) or "call" (a library function with its arguments;
    the stream is passed when run)."""

    def __init__(self, kind, name, fn=None, args=(), descs=()):
        from ._lib import GemmDesc
        self.kind, self.name, self.fn, self.args = kind, name, fn, args
        self.arr = (GemmDesc * max(1, len(descs)))(*[GemmDesc.from_buffer_copy(d) for d in descs])
        self.n = len(descs)

    def __call__(self, stream):
        if self.kind == "gemm":
            check(lib.pypde_dgemm_batched(ctypes.byref(self.arr), stream), self.name)
        elif self.kind == "group":
            check(lib.pypde_dgemm_grouped(ctypes.byref(self.arr), self.n, stream), self.name)
        else:
            check(self.fn(*self.args, stream), self.name)


class record_launches:
    """Context manager collecting the launches made inside it (a list of
    Recorded, plus any launch objects passed to record())."""

    def __enter__(self):
        self.launches = []
        _recorders.append(self.launches)
        return self.launches

    def __exit__(self, *exc):
        _recorders.remove(self.launches)
        return False


def record(launch):
    """Append a launch object (callable with a stream) to the active recorders."""
    for r in _recorders:
        r.append(launch)


def call(name, fn, *args):
    """fn(*args, stream) on the current stream, checked and recorded."""
    check(fn(*args, stream_ptr()), name)
    if _recorders:
        record(Recorded("call", name, fn=fn, args=args))


def _launch_desc(d):
    """Launch a GEMM descriptor with the autotuned kernel variant for its shape."""
    from . import autotune
    s = stream_ptr()
    autotune.tune_gemm_desc(d, lambda dd: check(lib.pypde_dgemm_batched(ctypes.byref(dd), s),
                                                "pypde_dgemm_batched"))
    check(lib.pypde_dgemm_batched(ctypes.byref(d), s), "pypde_dgemm_batched")
    if _recorders:
        record(Recorded("gemm", "gemm", descs=[d]))


def _desc(ta, tb, m, n, k, batch, alpha, beta, A, lda, sA, B, ldb, sB, D, ldd, sD, C, ldc, sC,
          alpha_vec=None):
    from ._lib import GemmDesc
    d = GemmDesc()
    d.transA, d.transB, d.m, d.n, d.k, d.batch = int(ta), int(tb), m, n, k, batch
    d.alpha, d.beta = float(alpha), float(beta)
    d.A, d.B, d.C = A, B, C
    d.D = D
    d.lda, d.ldb, d.ldd, d.ldc = lda, ldb, ldd, ldc
    d.iA, d.iB, d.iD, d.iC = _bi(sA), _bi(sB), _bi(sD), _bi(sC)
    d.alpha_vec = alpha_vec
    d.alpha_div = 1
    d.variant = 0
    return d


def matmul_axis(mat, x, axis=0, out=None, alpha=1.0, beta=0.0, add=None):
    """out = alpha * (mat applied along axis of x) + beta * add.

    axis 0: out[b] = mat @ x[b];  axis 1: out[b] = x[b] @ mat.T"""
    _check_axis(x, axis)
    m, k = mat.shape
    if _len(x, axis) != k:
        raise AssertionError("Shape mismatch: operator has %d columns, input axis has %d"
                             % (k, _len(x, axis)))
    if out is None:
        out = torch.empty(_out_shape(x, axis, m), dtype=DTYPE, device=x.device)
    x3, o3 = _as3(x), _as3(out)
    a3 = _as3(add) if add is not None else None
    B = x3.shape[0]
    matc = mat if mat.stride(1) == 1 else mat.contiguous()
    Dp = ptr(a3) if a3 is not None else None
    ldd = a3.stride(1) if a3 is not None else 0
    sd = a3.stride(0) if a3 is not None else 0
    if axis == 0:
        d = _desc(0, 0, m, x3.shape[2], k, B, alpha, beta, ptr(matc), matc.stride(0), 0,
                  ptr(x3), x3.stride(1), x3.stride(0), Dp, ldd, sd, ptr(o3), o3.stride(1),
                  o3.stride(0))
    else:
        d = _desc(0, 1, x3.shape[1], m, k, B, alpha, beta, ptr(x3), x3.stride(1), x3.stride(0),
                  ptr(matc), matc.stride(0), 0, Dp, ldd, sd, ptr(o3), o3.stride(1), o3.stride(0))
    _launch_desc(d)
    return out


def _mirror_sign(M, tol=1e-12):
    """s with M[n-1-j, k] = s (-1)^k M[j, k] for all j, k, else None."""
    par = (-1.0) ** np.arange(M.shape[1])
    scale = max(float(np.abs(M).max()), 1e-300)
    for sg in (1.0, -1.0):
        if np.abs(M[::-1] - sg * par * M).max() <= tol * scale:
            return sg
    return None


class MirrorOperator:
    """A dense operator M [n_out, n_in] applied along an axis (like matmul_axis)
    with the parity-folded GEMM when M is mirror-symmetric: on the output side,
    M[n_out-1-j, k] = s (-1)^k M[j, k] (backward transforms: half-size products
    over even and odd k, both mirrored outputs written by one launch), or on the
    input side with even n_in and n_out, M[k, n_in-1-j] = s (-1)^k M[k, j]
    (forward transforms: inputs folded X[j] +- X[n_in-1-j] in the loader, even
    and odd outputs as two batch entries).  Either halves the MFMA work; other
    operators fall back to the plain GEMM."""

    def __init__(self, M):
        M = np.asarray(M.detach().cpu().numpy() if isinstance(M, torch.Tensor) else M, dtype=np.float64)
        self.shape = M.shape
        self.M = to_dev(M)
        self.mode = 0
        n_out, n_in = M.shape
        so = _mirror_sign(M)
        si = _mirror_sign(M.T) if (n_in % 2 == 0 and n_out % 2 == 0) else None
        if si is not None:    # preferred: one accumulator set
            h = n_in // 2
            self.mode, self.sign = 2, si
            self.P = to_dev(np.stack([M[0::2, :h], M[1::2, :h]]))
        elif so is not None:
            h = (n_out + 1) // 2
            self.mode, self.sign = 1, so
            self.E = to_dev(np.ascontiguousarray(M[:h, 0::2]))
            self.O = to_dev(np.ascontiguousarray(M[:h, 1::2]))

    def apply(self, x, axis=0, out=None):
        _check_axis(x, axis)
        n_out, n_in = self.shape
        if _len(x, axis) != n_in:
            raise AssertionError("Shape mismatch: operator has %d columns, input axis has %d"
                                 % (n_in, _len(x, axis)))
        if self.mode == 0 or x.stride(-1) != 1:
            return matmul_axis(self.M, x, axis, out=out)
        if out is None:
            out = torch.empty(_out_shape(x, axis, n_out), dtype=DTYPE, device=x.device)
        x3, o3 = _as3(x), _as3(out)
        B = x3.shape[0]
        sx, so = x3.stride(0) if B > 1 else 0, o3.stride(0) if B > 1 else 0
        ke, ko = (n_in + 1) // 2, n_in // 2
        if self.mode == 1 and axis == 0:
            h = (n_out + 1) // 2
            nc, ld = x3.shape[2], x3.stride(1)
            d = _desc(0, 0, h, nc, ke, B, 1.0, 0.0, self.E.data_ptr(), ke, 0, x3.data_ptr(), 2 * ld,
                      sx, None, 0, 0, o3.data_ptr(), o3.stride(1), so)
            d.fold, d.A1, d.B1, d.lda1, d.ldb1, d.k1 = 1, self.O.data_ptr(), x3.data_ptr() + 8 * ld, ko, 2 * ld, ko
            d.fold_len, d.fold_sign = n_out, (1 if self.sign < 0 else 0)
        elif self.mode == 1:
            h = (n_out + 1) // 2
            nr, ld = x3.shape[1], x3.stride(1)
            d = _desc(0, 1, nr, h, ke, B, 1.0, 0.0, x3.data_ptr(), ld, sx, self.E.data_ptr(), ke, 0,
                      None, 0, 0, o3.data_ptr(), o3.stride(1), so)
            d.fold, d.A1, d.B1, d.lda1, d.ldb1, d.k1 = 2, x3.data_ptr() + 8, self.O.data_ptr(), ld, ko, ko
            d.ska, d.fold_len, d.fold_sign = 2, n_out, (1 if self.sign < 0 else 0)
        elif axis == 0:
            h, ho = n_in // 2, n_out // 2
            nc, ld, ldo = x3.shape[2], x3.stride(1), o3.stride(1)
            d = _desc(0, 0, ho, nc, h, 2 * B, 1.0, 0.0, self.P.data_ptr(), h, 0, x3.data_ptr(), ld, 0,
                      None, 0, 0, o3.data_ptr(), 2 * ldo, 0)
            d.iA, d.iB = _bi(ho * h, m1=2), _bi(sx, d1=2)
            d.iC = _bi(so, d1=2, s2=ldo, m2=2)
            d.fold, d.fold_len, d.mirror_sign, d.mirror_div = 4, n_in, self.sign, 1
        else:
            h, ho = n_in // 2, n_out // 2
            nr, ld = x3.shape[1], x3.stride(1)
            d = _desc(0, 1, nr, ho, h, 2 * B, 1.0, 0.0, x3.data_ptr(), ld, 0, self.P.data_ptr(), h, 0,
                      None, 0, 0, o3.data_ptr(), o3.stride(1), 0)
            d.iA, d.iB = _bi(sx, d1=2), _bi(ho * h, m1=2)
            d.iC = _bi(so, d1=2, s2=1, m2=2)
            d.csn = 2
            d.fold, d.fold_len, d.mirror_sign, d.mirror_div = 3, n_in, self.sign, 1
        _launch_desc(d)
        return out


def gemm(A, B, transA=False, transB=False, alpha=1.0, beta=0.0, D=None, out=None):
    """Plain (batched) row-major GEMM on 2-D or 3-D tensors."""
    A3 = A if A.dim() == 3 else A.unsqueeze(0)
    B3 = B if B.dim() == 3 else B.unsqueeze(0)
    m = A3.shape[2] if transA else A3.shape[1]
    k = A3.shape[1] if transA else A3.shape[2]
    kb = B3.shape[2] if transB else B3.shape[1]
    n = B3.shape[1] if transB else B3.shape[2]
    if k != kb:
        raise AssertionError("GEMM inner dimensions differ: %d vs %d" % (k, kb))
    batch = max(A3.shape[0], B3.shape[0])
    if out is None:
        shape = (batch, m, n) if (A.dim() == 3 or B.dim() == 3) else (m, n)
        out = torch.empty(shape, dtype=DTYPE, device=A.device)
    o3 = out if out.dim() == 3 else out.unsqueeze(0)
    D3 = None if D is None else (D if D.dim() == 3 else D.unsqueeze(0))
    d = _desc(transA, transB, m, n, k, batch, alpha, beta,
              ptr(A3), A3.stride(1), A3.stride(0) if A3.shape[0] > 1 else 0,
              ptr(B3), B3.stride(1), B3.stride(0) if B3.shape[0] > 1 else 0,
              ptr(D3) if D3 is not None else None, D3.stride(1) if D3 is not None else 0,
              (D3.stride(0) if D3.shape[0] > 1 else 0) if D3 is not None else 0,
              ptr(o3), o3.stride(1), o3.stride(0))
    _launch_desc(d)
    return out


def bgemm(A, B, transA=False, transB=False, alpha=1.0, alpha_vec=None, beta=0.0, D=None,
          out=None, defer=False):
    """Batched GEMM over a leading batch dimension with broadcasting of 2-D
    operands (one launch of pypde_dgemm_batched); alpha_vec [batch] scales
    entry e by alpha * alpha_vec[e].  defer: return (out, descriptor) without
    launching (for launch_grouped)."""

    def geo(T):
        if T.dim() == 2:
            return T, 1, 0
        return T, T.shape[0], T.stride(0)

    A3, ba, sa = geo(A)
    B3, bb, sb = geo(B)
    am, ak = (A3.shape[-1], A3.shape[-2]) if transA else (A3.shape[-2], A3.shape[-1])
    bk, bn = (B3.shape[-1], B3.shape[-2]) if transB else (B3.shape[-2], B3.shape[-1])
    if ak != bk:
        raise AssertionError("GEMM inner dimensions differ: %d vs %d" % (ak, bk))
    if A3.stride(-1) != 1 or B3.stride(-1) != 1:
        raise ValueError("operands need contiguous rows")
    batch = max(ba, bb)
    if out is None:
        out = torch.empty((batch, am, bn) if batch > 1 or A.dim() == 3 or B.dim() == 3 else (am, bn),
                          dtype=DTYPE, device=A.device)
    o3, bo, so = geo(out)
    if D is not None:
        D3, bd, sd = geo(D)
        Dp, ldd, sD = D3.data_ptr(), D3.stride(-2), sd if bd > 1 else 0
    else:
        Dp, ldd, sD = None, 0, 0
    d = _desc(transA, transB, am, bn, ak, batch, alpha, beta,
              A3.data_ptr(), A3.stride(-2), sa if ba > 1 else 0,
              B3.data_ptr(), B3.stride(-2), sb if bb > 1 else 0,
              Dp, ldd, sD, o3.data_ptr(), o3.stride(-2), so,
              alpha_vec.data_ptr() if alpha_vec is not None else None)
    if defer:
        return out, d
    _launch_desc(d)
    return out


class ChebTransform:
    """The square Chebyshev transform of N points along an axis ("bwd": values
    from coefficients, chebyshev.py:78-83; "fwd": coefficients from values with
    the mass weights, :67-76,143-149) in two kernel forms: the parity-folded
    MFMA product with the transform matrix (MirrorOperator), and the
    LDS-staged cosine sum on the vector units (pypde_cheb_dct, variants 1-3).
    apply() runs the form the autotuner timed fastest for the call's shape
    (the decision is stored with the others), so the choice is made per N,
    axis and batch from measurements on the device."""

    FORMS = ("mfma", 1, 2, 3)

    def __init__(self, kind, N, matrix):
        if kind not in ("bwd", "fwd"):
            raise ValueError("kind must be 'bwd' or 'fwd'")
        M = np.asarray(matrix.detach().cpu().numpy() if isinstance(matrix, torch.Tensor)
                       else matrix, dtype=np.float64)
        if M.shape != (N, N):
            raise ValueError("a square N x N transform matrix is required")
        self.kind, self.N, self.P = kind, N, N - 1
        self.mirror = MirrorOperator(M)
        P = self.P
        k = np.arange(N)
        sign = (-1.0) ** k
        if kind == "bwd":
            a, b = np.ones(N), sign
        else:
            w = np.full(N, 2.0)
            w[0] = w[-1] = 1.0
            a, b = w * sign / (2.0 * P), w
        m = np.arange(2 * P)
        ctab = np.cos(np.pi * m / P)
        # the factorised form must be the matrix
        C = ctab[(k[:, None] * k[None, :]) % (2 * P)]
        err = np.abs(a[:, None] * C * b[None, :] - M).max()
        if err > 1e-13 * np.abs(M).max():
            raise ValueError("matrix is not the %s Chebyshev transform (%.3e)" % (kind, err))
        self.a, self.b, self.ctab = to_dev(a), to_dev(b), to_dev(ctab)

    def _dct(self, x3, o3, axis, v):
        g = _geom(x3, o3, axis, self.N, self.N)
        check(lib.pypde_cheb_dct(ctypes.byref(g), self.P, ptr(self.a), ptr(self.b), ptr(self.ctab),
                                 v, stream_ptr()), "pypde_cheb_dct")

    def form(self, x, axis):
        """The kernel form used for x along axis (tuned on first use)."""
        from . import autotune
        key = ("cheb_transform", self.kind, self.N, axis, tuple(x.shape))
        known = autotune.cached(key, None)
        if known is not None:
            return known
        out = torch.empty(_out_shape(x, axis, self.N), dtype=DTYPE, device=x.device)
        x3, o3 = _as3(x), _as3(out)

        def run(f):
            if f == "mfma":
                self.mirror.apply(x, axis, out=out)
            else:
                self._dct(x3, o3, axis, f)
        return autotune.choose(key, run, self.FORMS, default="mfma")

    def apply(self, x, axis=0, out=None, form=None):
        _check_axis(x, axis)
        if _len(x, axis) != self.N:
            raise AssertionError("Shape mismatch: transform has %d points, input axis has %d"
                                 % (self.N, _len(x, axis)))
        f = self.form(x, axis) if form is None else form
        if f == "mfma" or x.stride(-1) != 1:
            return self.mirror.apply(x, axis, out=out)
        if out is None:
            out = torch.empty(_out_shape(x, axis, self.N), dtype=DTYPE, device=x.device)
        if out.data_ptr() == x.data_ptr():
            raise ValueError("the transform cannot run in place")
        self._dct(_as3(x), _as3(out), axis, f)
        return out


def launch_grouped(descs, tag):
    """Independent GEMM descriptors with the same transposes in one
    pypde_dgemm_grouped launch (at most four), the kernel variant tuned per
    group of shapes on first use."""
    from . import autotune
    from ._lib import GemmDesc
    n = len(descs)
    arr = (GemmDesc * n)(*descs)
    s = stream_ptr()

    def run(v):
        for i in range(n):
            arr[i].variant = v
        check(lib.pypde_dgemm_grouped(ctypes.byref(arr), n, s), "pypde_dgemm_grouped")
    key = ("ops_group", tag) + tuple(autotune.gemm_key(d) for d in descs)
    # a member whose output is also an input would be accumulated once per
    # timed launch: such a group is not timed (stored choice, else variant 0)
    if any(autotune.aliased(d) for d in descs):
        run(autotune.cached(key, default=0))
    else:
        run(autotune.choose(key, run, autotune.GEMM_CANDIDATES, default=0))
    if _recorders:
        record(Recorded("group", tag, descs=list(arr)))


# ---------------------------------------------------------------------------
def stencil_apply(s, x, axis=0, lpb=0):
    """u = S v along axis (Galerkin M -> Chebyshev M+2)."""
    _check_axis(x, axis)
    M = _len(x, axis)
    out = torch.empty(_out_shape(x, axis, M + 2), dtype=DTYPE, device=x.device)
    g = _geom(_as3(x), _as3(out), axis, M, M + 2, lpb)
    st = stream_ptr()

    def run(rows):
        g.lines_per_block = rows
        check(lib.pypde_stencil_apply(ctypes.byref(g), ptr(s), st), "pypde_stencil_apply")
    if lpb == 0:
        from . import autotune
        g.lines_per_block = autotune.choose(("stencil", axis, M, g.nlines, g.batch), run,
                                            autotune.rows_candidates(axis), default=1)
    run(g.lines_per_block)
    return out


def stencil_inverse(tabs, x, axis=0, lpb=0):
    """v = (S^T S)^-1 S^T u along axis (Chebyshev N -> Galerkin N-2).
    tabs = (s, a, b, w, den) device tensors."""
    _check_axis(x, axis)
    N = _len(x, axis)
    out = torch.empty(_out_shape(x, axis, N - 2), dtype=DTYPE, device=x.device)
    g = _geom(_as3(x), _as3(out), axis, N, N - 2, lpb)
    s, a, b, w, den = tabs
    check(lib.pypde_stencil_inverse(ctypes.byref(g), ptr(s), ptr(a), ptr(b), ptr(w), ptr(den),
                                    stream_ptr()), "pypde_stencil_inverse")
    return out


def cheb_deriv(x, axis=0, order=1, scale_pow=1.0, s=None, lpb=0):
    """Chebyshev derivative along axis (stencil first if s is given), / scale_pow."""
    _check_axis(x, axis)
    n_in = _len(x, axis)
    n = n_in + 2 if s is not None else n_in
    out = torch.empty(_out_shape(x, axis, n), dtype=DTYPE, device=x.device)
    g = _geom(_as3(x), _as3(out), axis, n_in, n, lpb)
    check(lib.pypde_cheb_deriv(ctypes.byref(g), ptr(s) if s is not None else None, int(order),
                               float(scale_pow), stream_ptr()), "pypde_cheb_deriv")
    return out


def banded_apply(offsets, vals, x, axis=0, len_out=None, lpb=0):
    """y = A x along axis for a banded A stored as diagonals vals[q, r] = A[r, r+offsets[q]]."""
    _check_axis(x, axis)
    n_in = _len(x, axis)
    n_out = vals.shape[1] if len_out is None else len_out
    out = torch.empty(_out_shape(x, axis, n_out), dtype=DTYPE, device=x.device)
    g = _geom(_as3(x), _as3(out), axis, n_in, n_out, lpb)
    offs = (ctypes.c_int * len(offsets))(*[int(o) for o in offsets])
    s = stream_ptr()

    def run(rows):
        g.lines_per_block = rows
        check(lib.pypde_banded_apply(ctypes.byref(g), len(offsets), offs, ptr(vals), s),
              "pypde_banded_apply")
    if lpb == 0:
        from . import autotune
        g.lines_per_block = autotune.choose(
            ("banded", axis, n_in, n_out, g.nlines, g.batch, tuple(int(o) for o in offsets)),
            run, autotune.rows_candidates(axis), default=1)
    run(g.lines_per_block)
    return out


def fdma_solve(tabs, x, axis=0, transpose=False, table_bs=0, lpb=0):
    """In-place pre-factored four-diagonal solve; tabs = (l, d, u1, u2)."""
    _check_axis(x, axis)
    n = _len(x, axis)
    x3 = _as3(x)
    g = _geom(x3, x3, axis, n, n, lpb)
    l, d, u1, u2 = tabs
    fn = lib.pypde_fdma_solve_t if transpose else lib.pypde_fdma_solve
    check(fn(ctypes.byref(g), ptr(l), ptr(d), ptr(u1), ptr(u2), int(table_bs), stream_ptr()),
          "pypde_fdma_solve")
    return x


def poisson_solve(tab, sing, x, axis=0, transpose=False, lpb=0):
    """In-place eigen-split Poisson line solves with per-line factor tables."""
    _check_axis(x, axis)
    n = _len(x, axis)
    x3 = _as3(x)
    g = _geom(x3, x3, axis, n, n, lpb)
    check(lib.pypde_poisson_solve(ctypes.byref(g), ptr(tab), ptr(sing), int(bool(transpose)),
                                  stream_ptr()), "pypde_poisson_solve")
    return x


def twodma_solve(d, u, x, axis=0, lpb=0):
    _check_axis(x, axis)
    n = _len(x, axis)
    x3 = _as3(x)
    g = _geom(x3, x3, axis, n, n, lpb)
    check(lib.pypde_twodma_solve(ctypes.byref(g), ptr(d), ptr(u), stream_ptr()),
          "pypde_twodma_solve")
    return x


def pdma_solve(tabs, x, axis=0, lpb=0):
    _check_axis(x, axis)
    n = _len(x, axis)
    x3 = _as3(x)
    g = _geom(x3, x3, axis, n, n, lpb)
    a, b, d, e, f = tabs
    check(lib.pypde_pdma_solve(ctypes.byref(g), ptr(a), ptr(b), ptr(d), ptr(e), ptr(f),
                               stream_ptr()), "pypde_pdma_solve")
    return x


def axpby(alpha, x, beta, y):
    """y = alpha*x + beta*y in place (same numel, contiguous)."""
    if x.numel() != y.numel():
        raise AssertionError("size mismatch")
    check(lib.pypde_axpby(y.numel(), float(alpha), ptr(x), float(beta), ptr(y), stream_ptr()),
          "pypde_axpby")
    return y


# ---------------------------------------------------------------------------
# register-resident recurrences (normalised tables, see include/pypde_amd.h)
# ---------------------------------------------------------------------------
class RecTables:
    """Device copy of a normalised recurrence (F, LM, R, MU, NU, pre tables).
    Per-line tables are kept in both layouts: [LM | R | MU | NU][n][nlines]
    for the block form and line-major [4][nlines][n] for the lane-chunk form."""

    def __init__(self, F=None, LM=None, R=None, MU=None, NU=None, pre=None, per_line=False,
                 entry_stride=0):
        if per_line:
            self.block = to_dev(np.stack([LM, R, MU, NU]))
            self.block_lm = self.block.transpose(1, 2).contiguous()
            self.t = dict(F=None, LM=self.block[0], R=self.block[1], MU=self.block[2],
                          NU=self.block[3], pre=None)
        else:
            self.t = {k: (to_dev(v) if v is not None else None)
                      for k, v in dict(F=F, LM=LM, R=R, MU=MU, NU=NU, pre=pre).items()}
        self.per_line = bool(per_line)
        self.entry_stride = int(entry_stride)
        self.has_pre = pre is not None

    def form(self):
        return tuple(self.t[k] is not None for k in ("F", "LM", "R", "NU", "pre"))

    def struct(self, variant=0, lines=0):
        """The C struct for kernel variant 1 (block, 64 lines per block), 2
        (lane-chunk), 3 (block, 32 lines per block) or 4 (as 3, 32 positions
        per chunk); 0 lets the library choose, except for per-line tables,
        whose layout follows the variant (lane-chunk up to 2048 lines)."""
        from ._lib import Recurrence
        r = Recurrence()
        tabs = self.t
        if self.per_line:
            if variant == 0:
                variant = 2 if lines <= 2048 else 1
            if variant == 2:
                tabs = dict(F=None, LM=self.block_lm[0], R=self.block_lm[1],
                            MU=self.block_lm[2], NU=self.block_lm[3], pre=None)
        for k, v in tabs.items():
            setattr(r, k, v.data_ptr() if v is not None else None)
        r.entry_stride = self.entry_stride
        r.per_line = (2 if variant == 2 else 1) if self.per_line else 0
        r.variant = variant
        return r


REC_CANDIDATES = (1, 2, 3, 4)


def recurrence_solve(tabs, x, axis=0, out=None, lpb=0, variant=None):
    """Banded line solve in normalised form along axis (in place when out is x
    and there is no pre-stencil).  variant None: the kernel form is autotuned
    per shape (timed out of place on scratch copies)."""
    _check_axis(x, axis)
    n_in = _len(x, axis)
    n = n_in - 2 if tabs.has_pre else n_in
    if out is None:
        out = x if not tabs.has_pre else torch.empty(_out_shape(x, axis, n), dtype=DTYPE,
                                                     device=x.device)
    g = _geom(_as3(x), _as3(out), axis, n_in, n, lpb)
    lines = g.nlines * g.batch
    s = stream_ptr()
    if variant is None:
        from . import autotune
        key = ("rec", axis, n_in, n, g.nlines, g.batch, tabs.form(), tabs.per_line)
        if autotune.enabled() and not autotune._capturing() and key not in autotune.cache():
            xs = _as3(x).clone()
            os_ = torch.empty(_out_shape(_as3(x), axis, n), dtype=DTYPE, device=x.device)
            gs = _geom(xs, os_, axis, n_in, n, lpb)

            def run(v):
                rs = tabs.struct(v, lines)
                check(lib.pypde_recurrence_solve(ctypes.byref(gs), ctypes.byref(rs), s),
                      "pypde_recurrence_solve")
            autotune.choose(key, run, REC_CANDIDATES, default=0)
        variant = autotune.cache().get(key, 0)
    r = tabs.struct(variant, lines)
    check(lib.pypde_recurrence_solve(ctypes.byref(g), ctypes.byref(r), s),
          "pypde_recurrence_solve")
    return out
