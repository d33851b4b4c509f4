"""ORACLE -- test infrastructure only.

A CPU restatement of preiter93/pypde's spectral-Galerkin RBC2D time step, used
as the checker for the HIP product path.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
this module; the product package never does.

It follows the reference operation by operation:

* third-party arithmetic is called exactly where the reference calls it
  (``scipy.fftpack.dctn(type=1)`` for the DCT-I, ``scipy.sparse`` products for
  stencils and plan matrices, ``numpy.linalg.eig/inv`` for the Poisson plan);
* the four f2py Fortran kernels run as the C restatement in
  ``oracle/fortran_kernels.c`` (checked bit-for-bit against the reference's own
  Fortran compiled into ``oracle/_ref``).

Pinning (what checks each part; see tests/test_oracle_*.py):
  transforms        closed-form cosine sums, numpy.polynomial.chebyshev
  stencils / maps   dense S, lstsq
  derivative        numpy.polynomial.chebyshev.chebder
  B (pseudo-inverse) closed form + (B D2)[rows 2:] = I
  banded solves     np.linalg.solve on the dense band; oracle/_ref bit-exact
  Helmholtz ADI     dense Kronecker (Ax (x) Ay) solve
  Poisson           dense Kronecker (Ax (x) Cy + Cx (x) Ay) solve, (0,0) gauge
  NS update         golden trajectories written by tests/golden/make_golden.py
  rounding floor    dense_transforms(): the same step with dense transforms
  Nusselt           conductive state gives Nu = 1 exactly
"""
import contextlib
import ctypes
import os

import numpy as np
import scipy.sparse as sp
from scipy.fftpack import dctn as _dctn

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_kernels.so")
if not os.path.exists(_LIB_PATH):
    raise ImportError(
        "oracle/liboracle_kernels.so missing: run `make -C oracle port` "
        "(or __graft_entry__.build())")
_lib = ctypes.CDLL(_LIB_PATH)
_dp = ctypes.POINTER(ctypes.c_double)
_i = ctypes.c_int
_lib.ora_diff_2d.argtypes = [_dp, _dp, _i, _i]
_lib.ora_tdma_2d.argtypes = [_dp, _dp, _dp, _dp, _i, _dp, _i, _i]
_lib.ora_fdma_2d.argtypes = [_dp, _dp, _dp, _dp, _dp, _i, _i, _i]
_lib.ora_fdma_1d.argtypes = [_dp, _dp, _dp, _dp, _dp, _i]
_lib.ora_init_fdma.argtypes = [_dp, _dp, _dp, _dp, _dp, _i]
_lib.ora_fdma_type2.argtypes = [_dp, _dp, _dp, _dp, _i, _i, _i, _i]
_lib.ora_twodma_2d.argtypes = [_dp, _dp, _dp, _i, _i, _i]


def _p(a):
    return a.ctypes.data_as(_dp)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# --------------------------------------------------------------------------
# Fortran kernels (C restatement), numpy-facing like the f2py modules
# --------------------------------------------------------------------------
def diff_2d(c):
    """differentiate_cheby.diff_2d: derivative recurrence along axis 0."""
    c = _c(c)
    two_d = c.ndim == 2
    c2 = c if two_d else c[:, None]
    c2 = _c(c2)
    dc = np.empty_like(c2)
    _lib.ora_diff_2d(_p(c2), _p(dc), c2.shape[0], c2.shape[1])
    return dc if two_d else dc[:, 0]


def tdma_solve(a, b, c, d, k):
    """tdma.solve_tdma_{1d,2d}: offset-k tridiagonal solve along axis 0."""
    d = _c(d)
    two_d = d.ndim == 2
    d2 = d if two_d else d[:, None]
    d2 = _c(d2)
    x = np.empty_like(d2)
    _lib.ora_tdma_2d(_p(_c(a)), _p(_c(b)), _p(_c(c)), _p(d2), int(k), _p(x),
                     d2.shape[0], d2.shape[1])
    return x if two_d else x[:, 0]


def fdma_solve(l, d, u1, u2, x, axis=0):
    """fdma.solve_fdma_{1d,2d}: in-place pre-factored four-diagonal solve."""
    xc = _c(x)
    if xc.ndim == 1:
        _lib.ora_fdma_1d(_p(_c(l)), _p(_c(d)), _p(_c(u1)), _p(_c(u2)), _p(xc), xc.shape[0])
    else:
        _lib.ora_fdma_2d(_p(_c(l)), _p(_c(d)), _p(_c(u1)), _p(_c(u2)), _p(xc), int(axis),
                         xc.shape[0], xc.shape[1])
    x[...] = xc
    return x


def init_fdma(A):
    """fdma.init_fdma: extract diagonals -2,0,2,4 and LU-factor them."""
    A = _c(A)
    n = A.shape[0]
    d, u1, u2, l = np.zeros(n), np.zeros(n - 2), np.zeros(max(n - 4, 1)), np.zeros(n - 2)
    _lib.ora_init_fdma(_p(A), _p(d), _p(u1), _p(u2), _p(l), n)
    return d, u1, u2[: n - 4], l


def fdma_type2(A, C, lam, x, axis, singular):
    """fdma.solve_fdma_type2: per-line dense (A + lam_i C) form + factor + solve."""
    xc = _c(x)
    _lib.ora_fdma_type2(_p(_c(A)), _p(_c(C)), _p(_c(lam)), _p(xc), int(axis),
                        int(bool(singular)), xc.shape[0], xc.shape[1])
    x[...] = xc
    return x


def twodma_solve(d, u, x, axis=0):
    """twodma.solve_twodma_{1d,2d}: in-place upper (0,+2) back substitution."""
    xc = _c(x)
    two_d = xc.ndim == 2
    x2 = xc if two_d else xc[:, None]
    x2 = _c(x2)
    _lib.ora_twodma_2d(_p(_c(d)), _p(_c(u)), _p(x2), int(axis) if two_d else 0,
                       x2.shape[0], x2.shape[1])
    x[...] = x2 if two_d else x2[:, 0]
    return x


def fdma_lu(l, d, u1, u2):
    """Plan_fdma.FDMA_LU (plans.py:226-234), the plan-time sweep (in place)."""
    n = d.shape[0]
    for i in range(2, n):
        l[i - 2] = l[i - 2] / d[i - 2]
        d[i] = d[i] - l[i - 2] * u1[i - 2]
        if i < n - 2:
            u1[i] = u1[i] - l[i - 2] * u2[i - 2]


def pdma_lu(a, b, d, e, f):
    """PDMA.PDMA_LU (pypde/solver/linalg/pdma.py:48-72), in place.
    a, b, d, e, f are the diagonals -4, -2, 0, 2, 4."""
    n = d.shape[0]
    m = e.shape[0]
    k = n - m
    for i in range(n - 2 * k):
        lam = b[i] / d[i]
        d[i + k] -= lam * e[i]
        e[i + k] -= lam * f[i]
        b[i] = lam
        lam = a[i] / d[i]
        b[i + k] -= lam * e[i]
        d[i + 2 * k] -= lam * f[i]
        a[i] = lam
    i = n - 4
    lam = b[i] / d[i]
    d[i + k] -= lam * e[i]
    b[i] = lam
    i = n - 3
    lam = b[i] / d[i]
    d[i + k] -= lam * e[i]
    b[i] = lam


def pdma_solve(a, b, d, e, f, h):
    """PDMA.PDMA_Solve (pdma.py:74-90) along axis 0 of h, in place."""
    n = d.shape[0]
    bc = h
    bc[2] -= b[0] * bc[0]
    bc[3] -= b[1] * bc[1]
    for k in range(4, n):
        bc[k] -= b[k - 2] * bc[k - 2] + a[k - 4] * bc[k - 4]
    bc[n - 1] /= d[n - 1]
    bc[n - 2] /= d[n - 2]
    bc[n - 3] = (bc[n - 3] - e[n - 3] * bc[n - 1]) / d[n - 3]
    bc[n - 4] = (bc[n - 4] - e[n - 4] * bc[n - 2]) / d[n - 4]
    for k in range(n - 5, -1, -1):
        bc[k] = (bc[k] - e[k] * bc[k + 2] - f[k] * bc[k + 4]) / d[k]
    return bc


def pdma_factor(A):
    """PDMA.init (pdma.py:32-46) on a dense pentadiagonal matrix."""
    A = np.asarray(A)
    d0 = np.diag(A, 0).copy()
    d1 = np.diag(A, 2).copy()
    d2 = np.diag(A, 4).copy()
    l1 = np.diag(A, -2).copy()
    l2 = np.diag(A, -4).copy()
    pdma_lu(l2, l1, d0, d1, d2)
    return l2, l1, d0, d1, d2


# --------------------------------------------------------------------------
# Grids and operator matrices
# --------------------------------------------------------------------------
def gauss_lobatto(n):
    """dmsuite.py:209-214: N=n+1 Chebyshev-Gauss-Lobatto points, increasing."""
    k = np.linspace(n, 0, n + 1)
    if n > 1:
        return np.sin(np.pi * (n - 2 * k) / (2 * n))
    return 0


def pseudoinverse_d2(N):
    """dmsuite.py:327-368 (deriv=2): Shen/Oh pseudo-inverse of the Chebyshev D2."""
    diag0 = np.zeros(N)
    diag0[2:-2] = np.array([-1 / (2 * (i ** 2 - 1)) for i in range(2, N - 2)])
    diag1 = np.zeros(N - 2)
    diag1[2:-2] = np.array([1 / (4 * i * (i + 1)) for i in range(2, N - 4)])
    diag2 = np.zeros(N - 2)
    diag2[:] = np.array([1 / (4 * i * (i - 1)) for i in range(2, N)])
    diag2[0] *= 2
    B = sp.diags([diag2, diag0, diag1], [-2, 0, 2]).toarray()
    # bases/utils.py:23-46 tosparse(): entries below 1e-12 are dropped
    B[np.abs(B) < 1e-12] = 0
    return B


def diff_mat_spectral(N, deriv):
    """dmsuite.py:218-268: dense spectral differentiation matrix (tests only)."""
    D = np.zeros((N, N))
    if deriv == 1:
        for n in range(N):
            for p in range(n + 1, N):
                if (p + n) % 2 != 0:
                    D[n, p] = p * 2
    elif deriv == 2:
        for n in range(N):
            for p in range(n + 2, N):
                if (p + n) % 2 == 0:
                    D[n, p] = p * (p ** 2 - n ** 2)
    else:
        raise NotImplementedError
    D[0, :] *= 0.5
    return D


def _tosparse_csc(A, tol=1e-12):
    A = A.copy()
    A[np.abs(A) < tol] = 0
    return sp.csc_matrix(A)


# --------------------------------------------------------------------------
# Bases (chebyshev.py)
# --------------------------------------------------------------------------
class Basis:
    """Chebyshev ("CH"), ChebDirichlet ("CD"), ChebNeumann ("CN"),
    ChebDirichletNeumann ("CDN") and the BC bases DirichletC ("DC") / NeumannC
    ("NC"); chebyshev.py:22-597."""

    def __init__(self, N, key, dealias=None):
        key = {"Chebyshev": "CH", "ChebDirichlet": "CD", "ChebNeumann": "CN",
               "ChebDirichletNeumann": "CDN", "DirichletC": "DC", "NeumannC": "NC"}.get(key, key)
        if key not in ("CH", "CD", "CN", "CDN", "DC", "NC"):
            raise ValueError("Key {:} not available.".format(key))
        self.key = self.id = key
        self.N = N
        self.x = gauss_lobatto(N - 1)
        self.is_galerkin = key != "CH"
        if key == "CH":
            self.M = N
        elif key in ("DC", "NC"):
            self.M = 2
        else:
            self.M = N - 2
        if self.is_galerkin:
            self.family = Basis(N, "CH")
            if key == "CD":
                self.bc = Basis(N, "DC")
            elif key == "CN":
                self.bc = Basis(N, "NC")
        if dealias is not None:
            self.dealias = Basis(int(N * dealias), key)

    # -- stencils (chebyshev.py:382-392, 426-436, 555-559, 586-590)
    def stencil(self):
        N, M = self.N, self.M
        if self.key == "CH":
            return np.eye(N)
        S = np.zeros((N, M))
        if self.key == "CD":
            for i in range(M):
                S[i, i], S[i + 2, i] = 1, -1
        elif self.key == "CN":
            for i in range(M):
                S[i, i], S[i + 2, i] = 1, -((i / (i + 2)) ** 2)
        elif self.key == "CDN":  # chebyshev.py:470-484
            k = np.arange(N)
            d1 = 4 * (k[:-1] + 1) / (2 * k[:-1] ** 2 + 6 * k[:-1] + 5)
            d2 = -(2 * k[:-2] ** 2 + 2 * k[:-2] + 1) / (2 * k[:-2] ** 2 + 6 * k[:-2] + 5)
            for i in range(M):
                S[i, i], S[i + 1, i], S[i + 2, i] = 1.0, d1[i], d2[i]
        elif self.key == "DC":
            S[0, 0], S[1, 0] = 0.5, -0.5
            S[0, 1], S[1, 1] = 0.5, 0.5
        elif self.key == "NC":
            S[0, 0], S[1, 0] = 0.5, -1 / 8
            S[0, 1], S[1, 1] = 0.5, 1 / 8
        return S

    @property
    def S(self):
        if not hasattr(self, "_S"):
            self._S = self.stencil()
        return self._S

    @property
    def S_sp(self):
        if not hasattr(self, "_S_sp"):
            self._S_sp = _tosparse_csc(self.S)
        return self._S_sp

    @property
    def ST_sp(self):
        if not hasattr(self, "_ST_sp"):
            self._ST_sp = _tosparse_csc(self.S.T)
        return self._ST_sp

    # -- Chebyshev transforms (chebyshev.py:67-93,143-149)
    def _cheb_forward(self, f):
        N = self.N
        sign = np.array([(-1) ** k for k in np.arange(N)])
        c = 0.5 * _dctn(f, type=1, axes=(0,)) / (N - 1)
        c = _prod(sign, c)
        m = np.array([1.0, *[2.0] * (N - 2), 1.0])
        return _prod(m, c)

    def _cheb_backward(self, c):
        N = self.N
        sign = np.array([(-1) ** k for k in np.arange(N)])
        f = _prod(sign, c)
        f[[0, -1]] = _prod(np.array([2, 2]), f[[0, -1]])
        return 0.5 * _dctn(f, type=1, axes=(0,))

    def forward_fft(self, f):
        if not self.is_galerkin:
            return self._cheb_forward(f)
        c = self.family._cheb_forward(f)
        return self.from_chebyshev(c)

    def backward_fft(self, c):
        if not self.is_galerkin:
            return self._cheb_backward(c)
        return self.family._cheb_backward(self.to_chebyshev(c))

    # -- Galerkin <-> Chebyshev (chebyshev.py:287-337,564-566,595-597)
    def to_chebyshev(self, vhat):
        assert vhat.shape[0] == self.M, "{} {}".format(vhat.shape[0], self.M)
        return self.S_sp @ vhat

    def from_chebyshev(self, uhat):
        assert uhat.shape[0] == self.N
        if self.key in ("DC", "NC"):
            return np.linalg.solve(self.S[:2, :2], uhat[:2])
        if self.key == "CDN":  # chebyshev.py:486-506: least squares
            return np.linalg.lstsq(self.S, uhat, rcond=None)[0]
        l2, d, u2 = self.stencil_inv_diags()
        return tdma_solve(l2, d, u2, self.ST_sp @ uhat, 2)

    def stencil_inv_diags(self):
        A = self.S.T @ self.S
        return tuple(np.diag(A, i) for i in (-2, 0, 2))

    # -- derivative (chebyshev.py:117-131, 347-363)
    def derivative(self, vhat, deriv, out_cheby=True):
        if not self.is_galerkin:
            if deriv == 0:
                return vhat
            df = vhat
            for _ in range(deriv):
                df = diff_2d(df)
            return df
        uhat = self.to_chebyshev(vhat)
        duhat = self.family.derivative(uhat, deriv)
        if out_cheby:
            return duhat
        return self.from_chebyshev(duhat)

    # -- pseudo-inverse (chebyshev.py:155-165)
    def B(self, deriv=2, discardrow=0):
        assert deriv == 2
        return pseudoinverse_d2(self.N)[discardrow:, :]

    def I(self, discardrow=0):
        return np.eye(self.N)[discardrow:, :]


@contextlib.contextmanager
def dense_transforms():
    """Within the block, Basis evaluates its Chebyshev transforms as dense
    products with the transform matrices (each the FFT-based transform applied
    to the identity) instead of scipy's DCT-I: the same mathematics with
    different rounding.  The distance between a trajectory computed this way
    and the ordinary one is the rounding floor of that trajectory, the scale
    on which two exact FP64 implementations of the reference differ."""
    fwd, bwd = Basis._cheb_forward, Basis._cheb_backward
    cache = {}

    def matrix(fn, basis, n):
        key = (fn.__name__, basis.N, n)
        if key not in cache:
            cache[key] = fn(basis, np.eye(n))
        return cache[key]

    Basis._cheb_forward = lambda self, f: matrix(fwd, self, f.shape[0]) @ f
    Basis._cheb_backward = lambda self, c: matrix(bwd, self, c.shape[0]) @ c
    try:
        yield
    finally:
        Basis._cheb_forward, Basis._cheb_backward = fwd, bwd


def _prod(a, b):
    """bases/utils.py:49-70 product(): multiply along the first dimension."""
    return a.reshape(a.shape + (1,) * (b.ndim - a.ndim)) * b


def zero_pad(array, target_length, axis=0):
    pad = target_length - array.shape[axis]
    if pad <= 0:
        return array
    npad = [(0, 0)] * array.ndim
    npad[axis] = (0, pad)
    return np.pad(array, pad_width=npad, mode="constant", constant_values=0)


def zero_unpad(array, target_length, axis=0):
    slc = [slice(None)] * array.ndim
    slc[axis] = slice(0, target_length)
    return array[tuple(slc)]


# --------------------------------------------------------------------------
# Fields (field.py, spectralspace.py)
# --------------------------------------------------------------------------
def _along(fn, a, axis):
    if axis == 0:
        return fn(a)
    return np.swapaxes(fn(np.swapaxes(a, axis, 0)), axis, 0)


class Field:
    def __init__(self, bases, _dealiased=False):
        self.xs = list(bases)
        self.ndim = len(self.xs)
        self.shape_physical = tuple(b.N for b in self.xs)
        self.shape_spectral = tuple(b.M for b in self.xs)
        self.v = np.zeros(self.shape_physical)
        self.vhat = np.zeros(self.shape_spectral)
        self.dealiased_space = _dealiased
        self.t = 0.0
        if not _dealiased and all(hasattr(b, "dealias") for b in self.xs):
            self.dealias = Field([b.dealias for b in self.xs], _dealiased=True)
            self.dealias.size_undealiased = [b.M for b in self.xs]

    @property
    def x(self):
        return self.xs[0].x

    @property
    def y(self):
        return self.xs[1].x

    @property
    def dx(self):
        x = self.x
        xm = np.zeros(x.size + 1)
        xm[0], xm[-1] = x[0], x[-1]
        xm[1:-1] = (x[1:] + x[:-1]) / 2.0
        return np.diff(xm)

    @property
    def dy(self):
        y = self.y
        ym = np.zeros(y.size + 1)
        ym[0], ym[-1] = y[0], y[-1]
        ym[1:-1] = (y[1:] + y[:-1]) / 2.0
        return np.diff(ym)

    def forward_fft(self, v, axis):
        return _along(self.xs[axis].forward_fft, v, axis)

    def backward_fft(self, vhat, axis):
        return _along(self.xs[axis].backward_fft, vhat, axis)

    def derivative(self, vhat, deriv, axis, out_cheby=True):
        return _along(lambda a: self.xs[axis].derivative(a, deriv, out_cheby), vhat, axis)

    def forward(self, v=None):
        vhat = self.v if v is None else v
        for axis in range(self.ndim):
            vhat = self.forward_fft(vhat, axis)
            if self.dealiased_space:
                vhat = zero_unpad(vhat, self.size_undealiased[axis], axis)
        if v is None:
            self.vhat = vhat
        else:
            return vhat

    def backward(self, vhat=None):
        v = self.vhat if vhat is None else vhat
        for axis in range(self.ndim):
            if self.dealiased_space:
                v = zero_pad(v, self.xs[axis].M, axis)
            v = self.backward_fft(v, axis)
        if vhat is None:
            self.v = v
        else:
            return v


class FieldBC(Field):
    """field.py:359-414: inhomogeneous part lifted from boundary values."""

    def __init__(self, bases, axis):
        xs = []
        for i, b in enumerate(bases):
            if i == axis:
                xs.append(b.bc if hasattr(b, "bc") else b)
            else:
                xs.append(b.family if hasattr(b, "family") else b)
        Field.__init__(self, xs, _dealiased=True)
        self.dealiased_space = False
        self.axis = axis

    def add_bc(self, bc):
        expected = list(self.shape_physical)
        expected[self.axis] = self.shape_spectral[self.axis]
        assert bc.shape == tuple(expected)
        self.v = self.backward_fft(bc, axis=self.axis)
        self.forward()


# --------------------------------------------------------------------------
# field_operations.py
# --------------------------------------------------------------------------
def grad(field, deriv, scale=None):
    dvhat = field.vhat
    for axis in range(field.ndim):
        dvhat = field.derivative(dvhat, deriv[axis], axis=axis)
        if scale is not None:
            dvhat = dvhat / scale[axis] ** deriv[axis]
    return dvhat


def galerkin_to_cheby(vhat, field):
    for axis in range(vhat.ndim):
        b = field.xs[axis]
        if b.is_galerkin:
            vhat = _along(b.to_chebyshev, vhat, axis)
    return vhat


def cheby_to_galerkin(uhat, field):
    for axis in range(uhat.ndim):
        b = field.xs[axis]
        if b.is_galerkin:
            uhat = _along(b.from_chebyshev, uhat, axis)
    return uhat


def conv_term(v_field, u, deriv, deriv_field, dealias, scale):
    vhat = grad(v_field, deriv, scale=scale)
    if dealias:
        dvdx = deriv_field.dealias.backward(vhat)
    else:
        dvdx = deriv_field.backward(vhat)
    return dvdx * u


def convective_term(v_field, ux, uz, deriv_field, add_bc=None, dealias=False, scale=None):
    conv = conv_term(v_field, ux, (1, 0), deriv_field, dealias, scale)
    conv = conv + conv_term(v_field, uz, (0, 1), deriv_field, dealias, scale)
    if add_bc is not None:
        conv = conv + add_bc
    if dealias:
        return deriv_field.dealias.forward(conv)
    return deriv_field.forward(conv)


def avg_x(f, dx):
    return np.sum(f * dx[:, None], axis=0) / np.sum(dx)


def avg_vol(f, dx, dy):
    favgx = np.sum(f * dx[:, None], axis=0) / np.sum(dx)
    return np.sum(favgx * dy) / np.sum(dy)


# --------------------------------------------------------------------------
# Plans (solver/plans.py, solverplan.py, templates/)
# --------------------------------------------------------------------------
class MatAxis:
    """PlanRHS / MatrixBase.dot (matrix.py:41-56): csr matrix along an axis."""

    def __init__(self, A, axis):
        self.A = sp.csr_matrix(np.asarray(A))
        self.dense = np.asarray(A)
        self.axis = axis

    def __call__(self, b):
        if self.axis == 0:
            return self.A @ b
        return np.swapaxes(self.A @ np.swapaxes(b, self.axis, 0), self.axis, 0)


class FdmaAxis:
    """Plan_fdma (plans.py:184-224): LU at plan time, Fortran sweep per call."""

    def __init__(self, A, axis):
        A = np.asarray(A)
        self.A = A
        self.l = np.diag(A, -2).copy()
        self.d = np.diag(A, 0).copy()
        self.u1 = np.diag(A, 2).copy()
        self.u2 = np.diag(A, 4).copy()
        fdma_lu(self.l, self.d, self.u1, self.u2)
        self.axis = axis

    def __call__(self, b):
        x = np.array(b, dtype=float, copy=True)
        return fdma_solve(self.l, self.d, self.u1, self.u2, x, self.axis)


class PoissonAxis:
    """Plan_Poisson.solve_fortran (plans.py:254-314)."""

    def __init__(self, A, alpha, C, axis, singular):
        self.A, self.alpha, self.C = np.asarray(A), np.asarray(alpha), np.asarray(C)
        self.axis, self.singular = axis, singular

    def __call__(self, b):
        x = np.array(b, dtype=float, copy=True)
        return fdma_type2(self.A, self.C, self.alpha, x, self.axis, self.singular)


class TwodmaAxis:
    """Plan_twodma (plans.py:137-181)."""

    def __init__(self, A, axis):
        A = np.asarray(A)
        self.d = np.diag(A, 0).copy()
        self.u = np.diag(A, 2).copy()
        self.axis = axis

    def __call__(self, b):
        x = np.array(b, dtype=float, copy=True)
        return twodma_solve(self.d, self.u, x, self.axis)


class SolverPlan:
    def __init__(self):
        self.rhs, self.old, self.lhs = [], [], []

    def solve_rhs(self, b):
        for p in self.rhs:
            b = p(b)
        return b

    def solve_old(self, b):
        for p in self.old:
            b = p(b)
        return b

    def solve_lhs(self, b):
        for p in self.lhs:
            b = p(b)
        return b


def hholtz_matrices(basis, lam, scale):
    """The four-diagonal A = B S - lam/s^2 I S and its RHS matrices."""
    S = basis.S_sp
    B = basis.family.B(2, 2)
    I = basis.family.I(2)
    A = B @ S - lam * (1.0 / scale ** 2.0) * I @ S
    return B, np.asarray(B @ S), np.asarray(A)


def solverplan_hholtz2d_adi(bases, lam, scale=(1, 1)):
    """templates/hholtz.py:42-93."""
    Bx, BSx, Ax = hholtz_matrices(bases[0], lam, scale[0])
    By, BSy, Ay = hholtz_matrices(bases[1], lam, scale[1])
    s = SolverPlan()
    s.rhs += [MatAxis(Bx, 0), MatAxis(By, 1)]
    s.old += [MatAxis(BSx, 0), MatAxis(BSy, 1)]
    s.lhs += [FdmaAxis(Ax, 0), FdmaAxis(Ay, 1)]
    s.mats = dict(Bx=Bx, By=By, BSx=BSx, BSy=BSy, Ax=Ax, Ay=Ay)
    return s


def eigdecomp(A):
    """solver/utils.py:13-29."""
    w, Q = np.linalg.eig(A)
    argsort = np.argsort(w)[::-1]
    w = w[argsort]
    Q = Q[:, argsort]
    Qi = np.linalg.inv(Q)
    return w, Q, Qi


def poisson_matrices(bases, singular=False, scale=(1, 1)):
    """templates/poisson.py:43-110 matrices."""
    bx, by = bases
    Sx, Bx, Ix = bx.S_sp, bx.family.B(2, 2), bx.family.I(2)
    Ax = np.asarray(Ix @ Sx * (1.0 / scale[0] ** 2.0))
    Cx = np.asarray(Bx @ Sx)
    Sy, By, Iy = by.S_sp, by.family.B(2, 2), by.family.I(2)
    Ay = np.asarray(Iy @ Sy * (1.0 / scale[1] ** 2.0))
    Cy = np.asarray(By @ Sy)
    CyI = np.linalg.inv(Cy)
    wy, Qy, QyI = eigdecomp(CyI @ Ay)
    if singular:
        wy[0] += 1e-20
    Hy = QyI @ CyI @ By
    return dict(Ax=Ax, Cx=Cx, Ay=Ay, Cy=Cy, Bx=Bx, By=By, wy=wy, Qy=Qy, QyI=QyI, Hy=Hy)


def solverplan_poisson2d(bases, singular=False, scale=(1, 1)):
    m = poisson_matrices(bases, singular, scale)
    s = SolverPlan()
    s.rhs += [MatAxis(m["Bx"], 0), MatAxis(m["Hy"], 1)]
    # the reference hard-codes singular=True in the LHS plan (poisson.py:106)
    s.lhs += [PoissonAxis(m["Ax"], m["wy"], m["Cx"], 0, True), MatAxis(m["Qy"], 1)]
    s.mats = m
    return s


# --------------------------------------------------------------------------
# navier/rbc2d_base.py + rbc2d.py
# --------------------------------------------------------------------------
def nu_fn(Ra, Pr, L):
    return np.sqrt(Pr / (Ra / L ** 3.0))


def kappa_fn(Ra, Pr, L):
    return np.sqrt(1 / Pr / (Ra / L ** 3.0))


def transfer_function(TL, TM, TR, x, k=0.01):
    arr = np.zeros(x.shape)
    L = x[-1] - x[0]
    for i in range(x.size):
        xs = x[i] * 2.0 / L
        if xs < 0:
            arr[i] = -k * xs / (k + xs + 1) * (TL - TM) + TM
        else:
            arr[i] = k * xs / (k - xs + 1) * (TR - TM) + TM
    return arr


class NavierStokes:
    """Oracle restatement of navier.rbc2d.NavierStokes (rbc2d.py:28-434)."""

    avail_cases = ["rbc", "linear", "zero"]

    def __init__(self, case="rbc", **kwargs):
        if case not in self.avail_cases:
            raise ValueError("Specified case is not available: ", self.avail_cases)
        if "Ra" in kwargs or "Pr" in kwargs:
            raise ValueError("Use small ra/pr!")
        self.case = case
        self.CONFIG = {"shape": (50, 50), "ra": 5e3, "pr": 1.0, "dt": 0.2, "ndim": 2,
                       "tsave": 0.1, "dealias": True, "integrator": "eu", "beta": 1.0,
                       "aspect": 1.0}
        self.CONFIG.update(**kwargs)
        self.__dict__.update(**self.CONFIG)
        self.time = 0.0
        self.set_nu_kappa()
        N0, N1 = self.shape
        self.deriv_field = Field([Basis(N0, "CH", 1.5), Basis(N1, "CH", 1.5)])
        self.x = self.deriv_field.x * self.scale[0]
        self.y = self.deriv_field.y * self.scale[1]
        side = "CN" if case == "rbc" else "CD"
        self.T = Field([Basis(N0, side, 1.5), Basis(N1, "CD", 1.5)])
        self.U = Field([Basis(N0, "CD", 1.5), Basis(N1, "CD", 1.5)])
        self.V = Field([Basis(N0, "CD", 1.5), Basis(N1, "CD", 1.5)])
        self.P = Field([Basis(N0, "CN"), Basis(N1, "CN")])
        self.pres = Field([Basis(N0, "CH"), Basis(N1, "CH")])
        self.setup_solver()
        if case == "zero":
            self.set_temp_fieldbc_zero()
        else:
            self.set_temp_fieldbc_linear()

    def set_nu_kappa(self):
        self.nu = nu_fn(self.ra, self.pr, L=1.0)
        self.kappa = kappa_fn(self.ra, self.pr, L=1.0)
        self.scale = (self.aspect * 0.5, 0.5)

    def grad(self, field, deriv):
        return grad(field, deriv, scale=self.scale)

    def set_coefficients(self):
        if self.integrator == "rk3":
            self.nstage = 3
            self.a = np.array([8.0 / 15.0, 2.0 / 15.0, 1.0 / 3.0])
            self.b = np.array([8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0])
            self.c = np.array([0, -17.0 / 60.0, -5.0 / 12.0])
        else:
            self.nstage = 1
            self.a, self.b, self.c = np.array([1.0]), np.array([1.0]), np.array([0])

    def setup_solver(self):
        self.set_coefficients()
        self.solver_U, self.solver_V, self.solver_T = [], [], []
        for rk in range(self.nstage):
            lam_u = self.dt * self.a[rk] * self.beta * self.nu
            lam_t = self.dt * self.a[rk] * self.beta * self.kappa
            self.solver_U.append(solverplan_hholtz2d_adi(self.U.xs, lam_u, self.scale))
            self.solver_V.append(solverplan_hholtz2d_adi(self.V.xs, lam_u, self.scale))
            self.solver_T.append(solverplan_hholtz2d_adi(self.T.xs, lam_t, self.scale))
        self.solver_P = solverplan_poisson2d(self.P.xs, singular=True, scale=self.scale)

    def _finish_bc(self):
        self.dTbcdz2 = self.grad(self.Tbc, (0, 2))
        vhat = self.grad(self.Tbc, (0, 1))
        if self.dealias:
            self.dTbcdz1 = self.deriv_field.dealias.backward(vhat)
        else:
            self.dTbcdz1 = self.deriv_field.backward(vhat)
        self.Tbc_cheby = galerkin_to_cheby(self.Tbc.vhat, self.Tbc)

    def set_temp_fieldbc_linear(self):
        bc = np.zeros((self.shape[0], 2))
        bc[:, 0], bc[:, 1] = 0.5, -0.5
        self.Tbc = FieldBC(self.T.xs, axis=1)
        self.Tbc.add_bc(bc)
        self._finish_bc()

    def set_temp_fieldbc_zero(self):
        bc = np.zeros((2, self.shape[1]))
        bc[0, :] = transfer_function(0.5, 0, -0.5, self.y, k=0.02)
        bc[1, :] = bc[0, :]
        self.Tbc = FieldBC(self.T.xs, axis=0)
        self.Tbc.add_bc(bc)
        self._finish_bc()

    def conv_term(self, field, ux, uz, add_bc=None):
        return convective_term(field, ux, uz, deriv_field=self.deriv_field, add_bc=add_bc,
                               dealias=self.dealias, scale=self.scale)

    def update_velocity(self, p, u, v, fac=1.0):
        dpdx = self.grad(p, (1, 0))
        dpdz = self.grad(p, (0, 1))
        u.vhat = u.vhat - cheby_to_galerkin(dpdx * fac, u)
        v.vhat = v.vhat - cheby_to_galerkin(dpdz * fac, v)

    def divergence_velocity(self, u, v):
        return self.grad(u, (1, 0)) + self.grad(v, (0, 1))

    def _add_explicit_diffusion(self, rhs, field, coef, stage):
        """rbc2d.py:269-283: rhs += f*d2x(field); rhs += f*d2y(field) (in this order)."""
        if self.beta != 1.0:
            f = self.dt * self.a[stage] * (1 - self.beta) * coef
            rhs = rhs + f * self.grad(field, (2, 0))
            rhs = rhs + f * self.grad(field, (0, 2))
        return rhs

    def update_U(self, stage):
        dpdx = self.grad(self.pres, (1, 0))
        rhs = -self.dt * self.a[stage] * dpdx
        rhs = rhs - self.dt * self.b[stage] * self.conv_term(self.U, self.ux, self.uz)
        if self.c[stage] != 0:
            rhs = rhs - self.dt * self.c[stage] * self.conv_term(self.U, self.ux_old, self.uz_old)
        rhs = self._add_explicit_diffusion(rhs, self.U, self.nu, stage)
        s = self.solver_U[stage]
        rhs = s.solve_rhs(rhs)
        rhs = rhs + s.solve_old(self.U.vhat)
        self.U.vhat = s.solve_lhs(rhs)

    def update_V(self, That, stage):
        dpdz = self.grad(self.pres, (0, 1))
        rhs = -self.dt * self.a[stage] * dpdz
        rhs = rhs - self.dt * self.b[stage] * self.conv_term(self.V, self.ux, self.uz)
        if self.c[stage] != 0:
            rhs = rhs - self.dt * self.c[stage] * self.conv_term(self.V, self.ux_old, self.uz_old)
        rhs = rhs + self.dt * self.a[stage] * That
        rhs = self._add_explicit_diffusion(rhs, self.V, self.nu, stage)
        s = self.solver_V[stage]
        rhs = s.solve_rhs(rhs)
        rhs = rhs + s.solve_old(self.V.vhat)
        self.V.vhat = s.solve_lhs(rhs)

    def update_T(self, stage):
        rhs = -self.dt * self.b[stage] * self.conv_term(
            self.T, self.ux, self.uz, add_bc=self.uz * self.dTbcdz1)
        if self.c[stage] != 0:
            rhs = rhs - self.dt * self.c[stage] * self.conv_term(
                self.T, self.ux_old, self.uz_old, add_bc=self.uz_old * self.dTbcdz1)
        rhs = rhs + self.dt * self.a[stage] * self.kappa * self.dTbcdz2
        rhs = self._add_explicit_diffusion(rhs, self.T, self.kappa, stage)
        s = self.solver_T[stage]
        rhs = s.solve_rhs(rhs)
        rhs = rhs + s.solve_old(self.T.vhat)
        self.T.vhat = s.solve_lhs(rhs)

    def update_P(self, div, singular=True):
        rhs = self.solver_P.solve_rhs(div)
        self.P.vhat = self.solver_P.solve_lhs(rhs)
        if singular:
            self.P.vhat[0, 0] = 0

    def update_pres(self, div, stage):
        self.pres.vhat = self.pres.vhat - 1.0 * self.nu * div * self.beta
        self.pres.vhat = self.pres.vhat + (
            1.0 / (self.dt * self.a[stage]) * galerkin_to_cheby(self.P.vhat, self.P))

    def update(self):
        self.ux_old, self.uz_old = 0, 0
        for rk in range(self.nstage):
            That = galerkin_to_cheby(self.T.vhat, self.T)
            That = That + self.Tbc_cheby
            if self.dealias:
                self.ux = self.U.dealias.backward(self.U.vhat)
                self.uz = self.V.dealias.backward(self.V.vhat)
            else:
                self.ux = self.U.backward(self.U.vhat)
                self.uz = self.V.backward(self.V.vhat)
            self.update_U(stage=rk)
            self.update_V(That, stage=rk)
            div = self.divergence_velocity(self.U, self.V)
            self.update_P(div)
            self.update_pres(div, stage=rk)
            self.update_velocity(self.P, self.U, self.V)
            self.update_T(stage=rk)
            self.ux_old, self.uz_old = self.ux, self.uz
        self.time += self.dt

    # -- diagnostics (rbc2d_base.py:196-203, 344-386)
    def eval_Nu(self):
        Lz = self.y[-1] - self.y[0]
        return eval_Nu(self.T, self.deriv_field, Lz=Lz, Tbc=self.Tbc), eval_Nuvol(
            self.T, self.V, self.kappa, self.deriv_field, Lz=Lz, Tbc=self.Tbc)

    def eval_KE(self):
        """Kinetic energy 0.5*avg_vol(u^2+v^2) on the N x N Gauss-Lobatto grid
        (the reference has no KE diagnostic; this definition is the harness's)."""
        u = self.U.backward(self.U.vhat)
        v = self.V.backward(self.V.vhat)
        f = self.deriv_field
        return 0.5 * avg_vol(u * u + v * v, f.dx, f.dy)

    def steady_residual(self):
        """rbc2d_base.py:321-341 steady_fun with dt=None: (Phi(X) - X)/dt."""
        X = np.concatenate([self.T.vhat.ravel(), self.U.vhat.ravel(), self.V.vhat.ravel()])
        self.update()
        Y = np.concatenate([self.T.vhat.ravel(), self.U.vhat.ravel(), self.V.vhat.ravel()])
        return (Y - X) / self.dt


class NavierStokesAdjoint:
    """Oracle restatement of navier.rbc2d_adj.NavierStokesAdjoint
    (rbc2d_adj.py:15-356): adjoint-descent iterations for the steady-state
    residual.  Each stage takes the residual R = (step(X) - X)/dt of an inner
    NavierStokes solver (whose pressure carries over between calls), smooths it
    by Poisson solves (nabla^2 A = R / nu, / kappa), and moves X along
    convective and residual terms of the adjoint fields, then projects the
    velocity as the forward step does."""

    def __init__(self, case="rbc", **kwargs):
        self.NS = NavierStokes(case=case, **kwargs)  # rbc2d_adj.py:44
        ns = self.NS
        for k in ("shape", "dt", "dealias", "nu", "kappa", "scale", "deriv_field", "a", "b",
                  "c", "nstage"):
            setattr(self, k, getattr(ns, k))
        N0, N1 = self.shape
        side = "CN" if case == "rbc" else "CD"

        def fld(kx, ky, dealias=1.5):
            return Field([Basis(N0, kx, dealias), Basis(N1, ky, dealias)])

        self.T, self.U, self.V = fld(side, "CD"), fld("CD", "CD"), fld("CD", "CD")
        self.P = Field([Basis(N0, "CN"), Basis(N1, "CN")])
        self.pres = Field([Basis(N0, "CH"), Basis(N1, "CH")])
        self.TA, self.UA, self.VA = fld(side, "CD"), fld("CD", "CD"), fld("CD", "CD")
        # plans (rbc2d_adj.py:131-144): the inner solver's pressure plan and
        # non-singular Poisson plans on each field's own bases for the smoothing
        self.solver_P = ns.solver_P
        self.nabla_U = solverplan_poisson2d(self.U.xs, singular=False, scale=self.scale)
        self.nabla_V = solverplan_poisson2d(self.V.xs, singular=False, scale=self.scale)
        self.nabla_T = solverplan_poisson2d(self.T.xs, singular=False, scale=self.scale)
        # conductive profile on the (dealiased) physical grid (rbc2d_adj.py:112-117)
        tbc = galerkin_to_cheby(ns.Tbc.vhat, ns.Tbc)
        df = self.deriv_field
        self.temp_bc = df.dealias.backward(tbc) if self.dealias else df.backward(tbc)

    def grad(self, field, deriv):
        return grad(field, deriv, scale=self.scale)

    def _phys(self, f):
        return f.dealias.backward(f.vhat) if self.dealias else f.backward(f.vhat)

    def _conv(self, field, u, deriv):
        return conv_term(field, u, deriv, self.deriv_field, self.dealias, self.scale)

    def _conv_adj(self, deriv, ux, uz, temp):
        """rbc2d_adj.py:162-187: d(UA) ux + d(VA) uz + d(TA) temp + d(TA) temp_bc,
        d = d/dx for deriv (1, 0) and d/dy for (0, 1), then the forward transform."""
        conv = self._conv(self.UA, ux, deriv)
        conv = conv + self._conv(self.VA, uz, deriv)
        conv = conv + self._conv(self.TA, temp, deriv)
        conv = conv + self._conv(self.TA, self.temp_bc, deriv)
        df = self.deriv_field
        return df.dealias.forward(conv) if self.dealias else df.forward(conv)

    def _residual(self):
        """rbc2d_adj.py:189-205."""
        ns = self.NS
        ns.U.vhat, ns.V.vhat, ns.T.vhat = self.U.vhat.copy(), self.V.vhat.copy(), self.T.vhat.copy()
        ns.update()
        ns.U.vhat = (ns.U.vhat - self.U.vhat) / self.dt
        ns.V.vhat = (ns.V.vhat - self.V.vhat) / self.dt
        ns.T.vhat = (ns.T.vhat - self.T.vhat) / self.dt
        for plan, src, dst, coef in ((self.nabla_U, ns.U, self.UA, self.nu),
                                     (self.nabla_V, ns.V, self.VA, self.nu),
                                     (self.nabla_T, ns.T, self.TA, self.kappa)):
            dst.vhat = plan.solve_lhs(plan.solve_rhs(galerkin_to_cheby(src.vhat, src))) / coef

    def _velocity_rhs(self, stage, deriv, own_A, resid):
        """rbc2d_adj.py:207-257 (U: deriv (1, 0); V: deriv (0, 1))."""
        a, b, c = self.a[stage], self.b[stage], self.c[stage]
        rhs = np.zeros(self.shape)
        rhs -= a * self.grad(self.pres, deriv)
        rhs += b * self.NS.conv_term(own_A, self.ux, self.uz)
        rhs += b * self._conv_adj(deriv, self.ux, self.uz, self.temp)
        if c != 0:
            rhs += c * self.NS.conv_term(own_A, self.ux_old, self.uz_old)
            rhs += c * self._conv_adj(deriv, self.ux_old, self.uz_old, self.temp_old)
        rhs += a * galerkin_to_cheby(resid.vhat, own_A)
        return rhs

    def update(self):
        """rbc2d_adj.py:321-356."""
        self.ux_old, self.uz_old, self.temp_old = 0, 0, 0
        for rk in range(self.nstage):
            a, b, c = self.a[rk], self.b[rk], self.c[rk]
            self.ux, self.uz, self.temp = self._phys(self.U), self._phys(self.V), self._phys(self.T)
            self._residual()
            ns = self.NS
            rhs = self._velocity_rhs(rk, (1, 0), self.UA, ns.U)
            self.U.vhat = self.U.vhat + self.dt * cheby_to_galerkin(rhs, self.U)
            rhs = self._velocity_rhs(rk, (0, 1), self.VA, ns.V)
            self.V.vhat = self.V.vhat + self.dt * cheby_to_galerkin(rhs, self.V)
            # projection (rbc2d_adj.py:283-315)
            div = ns.divergence_velocity(self.U, self.V)
            self.P.vhat = self.solver_P.solve_lhs(self.solver_P.solve_rhs(div))
            self.P.vhat[0, 0] = 0
            self.pres.vhat = self.pres.vhat + galerkin_to_cheby(self.P.vhat, self.P) / (self.dt * a)
            dpdx, dpdz = self.grad(self.P, (1, 0)), self.grad(self.P, (0, 1))
            self.U.vhat = self.U.vhat - cheby_to_galerkin(dpdx * 1.0, self.U)
            self.V.vhat = self.V.vhat - cheby_to_galerkin(dpdz * 1.0, self.V)
            # temperature (rbc2d_adj.py:259-281)
            rhs = np.zeros(self.shape)
            rhs += b * ns.conv_term(self.TA, self.ux, self.uz)
            if c != 0:
                rhs += c * ns.conv_term(self.TA, self.ux_old, self.uz_old)
            rhs += a * galerkin_to_cheby(ns.T.vhat, self.T)
            rhs += a * galerkin_to_cheby(self.VA.vhat, self.VA)
            self.T.vhat = self.T.vhat + self.dt * cheby_to_galerkin(rhs, self.T)
            self.ux_old, self.uz_old, self.temp_old = self.ux, self.uz, self.temp


def eval_Nu(T, field, Lz=1.0, Tbc=None):
    Tv = T.backward(T.vhat)
    if Tbc is not None:
        Tv = Tv + Tbc.v
    That = field.forward(Tv)
    scale = Lz / 2.0
    dThat = field.derivative(That, 1, axis=1) / scale
    dT = field.backward(dThat)
    dTavg = avg_x(dT, field.dx)
    Nu_bot = -dTavg[0] * Lz
    Nu_top = -dTavg[-1] * Lz
    return (Nu_bot + Nu_top) / 2.0


def eval_Nuvol(T, V, kappa, field, Lz=1.0, Tbc=None):
    Tv = T.backward(T.vhat)
    if Tbc is not None:
        Tv = Tv + Tbc.v
    Vv = V.backward(V.vhat)
    That = field.forward(Tv)
    scale = Lz / 2.0
    dThat = field.derivative(That, 1, axis=1) / scale
    dT = field.backward(dThat)
    Nuvol = (Tv * Vv / kappa - dT) * Lz
    return avg_vol(Nuvol, field.dx, field.dy)


def seeded_initial_state(ns, seed=0, amplitude=1e-3):
    """The harness's seeded initial condition: uniform noise of the given
    amplitude on the N x N Gauss-Lobatto grid assigned to T.v (the homogeneous
    part; the conductive profile lives in Tbc), then T.forward(); U=V=P=pres=0."""
    rng = np.random.default_rng(seed)
    ns.T.v = rng.uniform(-amplitude, amplitude, size=ns.T.shape_physical)
    ns.T.vhat = ns.T.forward(ns.T.v)
    return ns
