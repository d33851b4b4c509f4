"""Spectral bases on the GPU: Chebyshev, ChebDirichlet, ChebNeumann,
ChebDirichletNeumann and the boundary bases DirichletC / NeumannC.

Drop-in for pypde.bases (pypde/bases/spectralbase.py:12-45,
pypde/bases/chebyshev.py:22-597): same constructor (`Base(N, key, dealias=)`),
same attributes (N, M, x, id, family, bc, dealias, S, ST, S_sp, ST_sp) and
methods (forward_fft, backward_fft, to_chebyshev, from_chebyshev, derivative,
B, I).  Arrays are float64 device tensors (numpy input is accepted and moved to
the device); each method acts along axis 0 of its input, or along `axis`.

Operator tables are built once per basis on the host in float64 and cached on
the device.  Dense transforms run as MFMA GEMMs against the collocation
matrices, stencil maps and the offset-2 tridiagonal solve as line kernels.
"""
import numpy as np
import scipy.sparse as sp
import torch

from . import ops
from .ops import DTYPE


def gauss_lobatto(n):
    """N = n+1 Chebyshev-Gauss-Lobatto points, increasing: x_j = -cos(pi j / n)
    (pypde/bases/dmsuite.py:209-214 computes the same points as a sine)."""
    k = np.linspace(n, 0, n + 1)
    if n > 1:
        return np.sin(np.pi * (n - 2 * k) / (2 * n))
    return 0


def _cos_table(nrow, ncol, L):
    """cos(pi * j * k / (L-1)) for j < nrow, k < ncol, with exact argument reduction."""
    j = np.arange(nrow)[:, None]
    k = np.arange(ncol)[None, :]
    a = (j * k) % (2 * (L - 1))
    return np.cos(np.pi * a / (L - 1))


def cheb_backward_matrix(L, N):
    """Values at the L Gauss-Lobatto points of a Chebyshev series with N <= L
    coefficients: B[j, k] = T_k(x_j) = (-1)^k cos(pi j k / (L-1)).
    Equal to chebyshev.py:78-83 (sign, doubled end terms, 0.5*DCT-I) with the
    zero padding of utils.py:88-109 folded in."""
    sign = (-1.0) ** np.arange(N)
    return _cos_table(L, N, L) * sign[None, :]


def cheb_forward_matrix(L, N):
    """First N <= L Chebyshev coefficients of L samples at the Gauss-Lobatto
    points (chebyshev.py:67-76 and solve_mass :143-149, then zero_unpad):
    c_k = w_k (-1)^k / (2(L-1)) * [u_0 + (-1)^k u_{L-1} + 2 sum_j u_j cos(pi j k/(L-1))]."""
    C = _cos_table(N, L, L)  # [k, j]
    C[:, 1:-1] *= 2.0
    w = np.full(N, 2.0)
    w[0] = 1.0
    if N == L:
        w[-1] = 1.0
    sign = (-1.0) ** np.arange(N)
    return C * (w * sign / (2.0 * (L - 1)))[:, None]


def pseudoinverse_d2(N):
    """Shen/Oh pseudo-inverse of the Chebyshev second-derivative matrix
    (pypde/bases/dmsuite.py:327-368 with deriv=2, entries below 1e-12 dropped as
    bases/utils.py:23-46 does).  B[i,i-2] = 1/(4i(i-1)) (B[2,0] = 1/4),
    B[i,i] = -1/(2(i^2-1)) for 2 <= i <= N-3, B[i,i+2] = 1/(4i(i+1)) for 2 <= i <= N-5."""
    B = np.zeros((N, N))
    for i in range(2, N):
        B[i, i - 2] = 1.0 / (4 * i * (i - 1))
    B[2, 0] *= 2
    for i in range(2, N - 2):
        B[i, i] = -1.0 / (2 * (i ** 2 - 1))
    for i in range(2, N - 4):
        B[i, i + 2] = 1.0 / (4 * i * (i + 1))
    B[np.abs(B) < 1e-12] = 0
    return B


def _tosparse(A, tol=1e-12, fmt="csc"):
    A = np.array(A, dtype=float)
    A[np.abs(A) < tol] = 0
    return sp.csc_matrix(A) if fmt == "csc" else sp.csr_matrix(A)


class _DeviceCache:
    """Per-basis cache of device tensors built from host numpy arrays."""

    def __init__(self):
        self._t = {}

    def get(self, key, builder):
        dev = torch.cuda.current_device()
        k = (key, dev)
        if k not in self._t:
            self._t[k] = ops.to_dev(builder())
        return self._t[k]


class MetaBase:
    """Common part of all bases (pypde/bases/spectralbase.py:48-247)."""

    def __init__(self, N, x, dealias=None):
        self._N = N
        self._x = x
        self.name = self.__class__.__name__
        self.id = None
        self._dev = _DeviceCache()
        self._mops = {}
        if dealias is not None:
            self.create_dealiased_base(self.N * dealias)

    @property
    def x(self):
        return self._x

    @property
    def N(self):
        return self._N

    @property
    def M(self):
        return len(range(*self.slice().indices(self.N)))

    def slice(self):
        return slice(0, self.N)

    def _mirror(self, key, mat):
        """Cached ops.MirrorOperator of a device operator table."""
        op = self._mops.get(key)
        if op is None:
            op = self._mops[key] = ops.MirrorOperator(mat)
        return op

    def create_dealiased_base(self, size):
        self.dealias = self.__class__(int(size), dealias=None)

    # -- host matrices
    @property
    def S(self):
        if hasattr(self, "stencil"):
            return self.stencil()
        return np.eye(self.N)

    @property
    def ST(self):
        if hasattr(self, "stencil"):
            return self.stencil(transpose=True)
        return np.eye(self.N)

    @property
    def S_sp(self):
        return _tosparse(self.S)

    @property
    def ST_sp(self):
        return _tosparse(self.ST)


class Chebyshev(MetaBase):
    """T_k on the Gauss-Lobatto grid (pypde/bases/chebyshev.py:22-165)."""

    def __init__(self, N, dealias=None):
        MetaBase.__init__(self, N, gauss_lobatto(N - 1), dealias)
        self.id = "CH"
        self.family_id = "CH"

    def get_basis(self, i=0, x=None):
        x = self.x if x is None else x
        return np.cos(i * np.arccos(x))

    # -- device operator tables
    def backward_matrix(self, n_coef=None):
        n = self.N if n_coef is None else n_coef
        return self._dev.get(("bwd", n), lambda: cheb_backward_matrix(self.N, n))

    def forward_matrix(self, n_coef=None):
        n = self.N if n_coef is None else n_coef
        return self._dev.get(("fwd", n), lambda: cheb_forward_matrix(self.N, n))

    def forward_fft(self, f, mass=True, axis=0, n_coef=None):
        """Chebyshev coefficients of samples f at the grid (first n_coef of them)."""
        if not mass:
            raise NotImplementedError("forward_fft(mass=False) is not provided")
        f = ops.to_dev(f)
        if n_coef is None or n_coef == self.N:
            return self._transform("fwd").apply(f, axis)
        return self._mirror(("fwd", n_coef), self.forward_matrix(n_coef)).apply(f, axis)

    def backward_fft(self, c, axis=0):
        """Values at the grid of a series with c.shape[axis] <= N coefficients."""
        c = ops.to_dev(c)
        n = ops._len(c, axis)
        if n > self.N:
            raise AssertionError("Shape mismatch (%d) (%d)" % (n, self.N))
        if n == self.N:
            return self._transform("bwd").apply(c, axis)
        return self._mirror(("bwd", n), self.backward_matrix(n)).apply(c, axis)

    def _transform(self, kind):
        """Cached ops.ChebTransform (square transform, kernel form tuned per shape)."""
        op = self._mops.get(("cheb", kind))
        if op is None:
            mat = self.forward_matrix() if kind == "fwd" else self.backward_matrix()
            op = self._mops[("cheb", kind)] = ops.ChebTransform(kind, self.N, mat)
        return op

    def derivative(self, fhat, deriv, out_cheby=True, axis=0, scale_pow=1.0):
        fhat = ops.to_dev(fhat)
        if deriv == 0:
            return fhat if scale_pow == 1.0 else fhat / scale_pow
        return ops.cheb_deriv(fhat, axis, deriv, scale_pow)

    def B(self, deriv, discardrow=0):
        if deriv > 2:
            raise ValueError("deriv>2 not supported")
        if deriv != 2:
            raise NotImplementedError("only the D2 pseudo-inverse is provided")
        return pseudoinverse_d2(self.N)[discardrow:, :]

    def I(self, discardrow=0):
        return np.eye(self.N)[discardrow:, :]


class GalerkinChebyshev(MetaBase):
    """Composite bases phi_k = S T (pypde/bases/chebyshev.py:168-363)."""

    def __init__(self, N, dealias=None):
        MetaBase.__init__(self, N, gauss_lobatto(N - 1), dealias)
        self._bc = None
        self.family_id = "CH"
        self.family = Chebyshev(self.N)

    def slice(self):
        return slice(0, self.N - 2)

    def _stencil(self):
        raise NotImplementedError

    def stencil(self, transpose=False):
        S = self._stencil()
        return S.T if transpose else S

    def get_basis(self, i=0, x=None):
        from numpy.polynomial import chebyshev
        x = self.x if x is None else np.atleast_1d(x)
        if i >= self.M:
            raise ValueError("basis not known for i={:4d}".format(i))
        return chebyshev.Chebyshev(self.stencil()[:, i])(x)

    # -- device tables
    def stencil_vector(self):
        """s with S[i,i]=1, S[i+2,i]=s[i] (sub-diagonal -2 of the stencil)."""
        return self._dev.get("s", lambda: np.array([self._stencil()[i + 2, i] for i in range(self.M)]))

    def stencil_inv_tables(self):
        """(s, a, b, w, den) for (S^T S)^-1 S^T: a, b the -2 / 0 diagonals of
        S^T S, w and den the plan-time factors of the offset-2 sweep
        (tdma.f90:82-97), computed in float64 exactly as the Fortran does."""
        def build():
            S = self._stencil()
            A = S.T @ S
            a, b, c = np.diag(A, -2), np.diag(A, 0), np.diag(A, 2)
            n, k = b.shape[0], 2
            w = np.zeros(n - k)
            for i in range(n - k):
                w[i] = c[i] / b[i] if i < k else c[i] / (b[i] - a[i - k] * w[i - k])
            den = np.array(b, copy=True)
            for i in range(k, n):
                den[i] = b[i] - a[i - k] * w[i - k]
            return a, b, w, den
        if not hasattr(self, "_inv_tabs_host"):
            a, b, w, den = build()
            self._inv_tabs_host = (a, b, w, den)
        a, b, w, den = self._inv_tabs_host
        return (self.stencil_vector(), self._dev.get("ta", lambda: a), self._dev.get("tb", lambda: b),
                self._dev.get("tw", lambda: w), self._dev.get("tden", lambda: den))

    def backward_matrix(self, n_coef=None):
        """Grid values from the first n_coef Galerkin coefficients: T S[:, :n_coef]."""
        n = self.M if n_coef is None else n_coef

        def build():
            S = self._stencil()[:, :n]
            nz = n + 2  # rows of S that can be nonzero
            return cheb_backward_matrix(self.N, min(nz, self.N)) @ S[: min(nz, self.N)]
        return self._dev.get(("gbwd", n), build)

    def forward_fft(self, f, bc=None, axis=0):
        f = ops.to_dev(f)
        if bc is not None:
            f = f - self.eval_inhomogeneous(bc, axis=axis)
        c = self.family.forward_fft(f, axis=axis)
        return self.from_chebyshev(c, axis=axis)

    def backward_fft(self, c, bc=None, axis=0):
        c = ops.to_dev(c)
        n = ops._len(c, axis)
        if n > self.M:
            raise AssertionError("Shape mismatch (%d) (%d)" % (n, self.M))
        v = self._mirror(("gbwd", n), self.backward_matrix(n)).apply(c, axis)
        if bc is not None:
            v = v + self.bc.backward_fft(bc, axis=axis)
        return v

    def to_chebyshev(self, vhat, axis=0):
        vhat = ops.to_dev(vhat)
        assert ops._len(vhat, axis) == self.M, "{} {}".format(ops._len(vhat, axis), self.M)
        return ops.stencil_apply(self.stencil_vector(), vhat, axis)

    def projection_tables(self):
        """Normalised recurrence of (S^T S)^-1 S^T for the register-resident kernel."""
        if not hasattr(self, "_proj"):
            from .solver import rec_from_tdma
            self.stencil_inv_tables()
            a, b, w, den = self._inv_tabs_host
            s = np.array([self._stencil()[i + 2, i] for i in range(self.M)])
            pre, F, LM, MU = rec_from_tdma(s, a, b, w, den)
            self._proj = ops.RecTables(pre=pre, F=F, LM=LM, MU=MU)
        return self._proj

    def from_chebyshev(self, uhat, axis=0, exact=False):
        uhat = ops.to_dev(uhat)
        assert ops._len(uhat, axis) == self.N, "Shape mismatch ({:3}) ({:3})".format(
            ops._len(uhat, axis), self.N)
        if exact or self.M < 3:
            return ops.stencil_inverse(self.stencil_inv_tables(), uhat, axis)
        return ops.recurrence_solve(self.projection_tables(), uhat, axis)

    def eval_inhomogeneous(self, bchat, axis=0):
        return self.bc.backward_fft(bchat, axis=axis)

    def derivative(self, vhat, deriv, out_cheby=True, axis=0, scale_pow=1.0):
        vhat = ops.to_dev(vhat)
        if deriv == 0:
            u = self.to_chebyshev(vhat, axis)
            duhat = u if scale_pow == 1.0 else u / scale_pow
        else:
            duhat = ops.cheb_deriv(vhat, axis, deriv, scale_pow, s=self.stencil_vector())
        if out_cheby:
            return duhat
        return self.from_chebyshev(duhat, axis)


class ChebDirichlet(GalerkinChebyshev):
    """phi_k = T_k - T_{k+2} (pypde/bases/chebyshev.py:366-392)."""

    def __init__(self, N, dealias=None):
        GalerkinChebyshev.__init__(self, N, dealias)
        self.id = "CD"
        self.bc = DirichletC(N)

    def _stencil(self):
        S = np.zeros((self.N, self.M))
        for i in range(self.M):
            S[i, i], S[i + 2, i] = 1, -1
        return S


class ChebNeumann(GalerkinChebyshev):
    """phi_k = T_k - k^2/(k+2)^2 T_{k+2} (pypde/bases/chebyshev.py:410-436)."""

    def __init__(self, N, dealias=None):
        GalerkinChebyshev.__init__(self, N, dealias)
        self.id = "CN"
        self.bc = NeumannC(N)

    def _stencil(self):
        S = np.zeros((self.N, self.M))
        for i in range(self.M):
            S[i, i], S[i + 2, i] = 1, -((i / (i + 2)) ** 2)
        return S


class ChebDirichletNeumann(GalerkinChebyshev):
    """u(-1) = u'(1) = 0: phi_k = T_k + d1_k T_{k+1} + d2_k T_{k+2}
    (pypde/bases/chebyshev.py:454-530).

    The stencil has a -1 diagonal, so to_chebyshev is a three-diagonal banded
    apply, and from_chebyshev the least-squares map the reference computes per
    call with lstsq(S, u): pinv(S), an M x N table applied as one GEMM."""

    def __init__(self, N, dealias=None):
        GalerkinChebyshev.__init__(self, N, dealias)
        self.id = "CDN"
        self.bc = None

    def _stencil(self):
        S = np.zeros((self.N, self.M))
        k = np.arange(self.N, dtype=float)
        den = 2 * k ** 2 + 6 * k + 5
        d1 = 4 * (k + 1) / den
        d2 = -(2 * k ** 2 + 2 * k + 1) / den
        for i in range(self.M):
            S[i, i], S[i + 1, i], S[i + 2, i] = 1.0, d1[i], d2[i]
        return S

    def _band(self):
        """Diagonals of S (N x M) for the banded apply: vals[q, r] = S[r, r + off[q]]."""
        def build():
            S = self._stencil()
            vals = np.zeros((3, self.N))
            for q, off in enumerate((-2, -1, 0)):
                for r in range(self.N):
                    if 0 <= r + off < self.M:
                        vals[q, r] = S[r, r + off]
            return vals
        return self._dev.get("band", build)

    def to_chebyshev(self, vhat, axis=0):
        vhat = ops.to_dev(vhat)
        assert ops._len(vhat, axis) == self.M, "{} {}".format(ops._len(vhat, axis), self.M)
        return ops.banded_apply((-2, -1, 0), self._band(), vhat, axis, len_out=self.N)

    def from_chebyshev(self, uhat, axis=0, exact=False):
        uhat = ops.to_dev(uhat)
        assert ops._len(uhat, axis) == self.N, "Shape mismatch ({:3}) ({:3})".format(
            ops._len(uhat, axis), self.N)
        P = self._dev.get("pinv", lambda: np.linalg.pinv(self._stencil()))
        return ops.matmul_axis(P, uhat, axis)

    def derivative(self, vhat, deriv, out_cheby=True, axis=0, scale_pow=1.0):
        u = self.to_chebyshev(vhat, axis)
        duhat = self.family.derivative(u, deriv, axis=axis, scale_pow=scale_pow)
        if out_cheby:
            return duhat
        return self.from_chebyshev(duhat, axis)


class _BCBase(GalerkinChebyshev):
    """Two-function bases carrying boundary values (chebyshev.py:538-597)."""

    def slice(self):
        return slice(0, 2)

    def backward_matrix(self, n_coef=None):
        n = 2 if n_coef is None else n_coef
        return self._dev.get(("gbwd", n), lambda: cheb_backward_matrix(self.N, 2) @ self._stencil()[:2, :n])

    def to_chebyshev(self, vhat, axis=0):
        vhat = ops.to_dev(vhat)
        assert ops._len(vhat, axis) == 2
        return ops.matmul_axis(self._dev.get("S", self._stencil), vhat, axis)

    def derivative(self, vhat, deriv, out_cheby=True, axis=0, scale_pow=1.0):
        u = self.to_chebyshev(vhat, axis)
        duhat = self.family.derivative(u, deriv, axis=axis, scale_pow=scale_pow)
        if out_cheby:
            return duhat
        return self.from_chebyshev(duhat, axis)

    def from_chebyshev(self, c, axis=0):
        c = ops.to_dev(c)
        Sinv = self._dev.get("Sinv2", lambda: np.hstack([np.linalg.inv(self._stencil()[:2, :2]),
                                                          np.zeros((2, self.N - 2))]))
        return ops.matmul_axis(Sinv, c, axis)


class DirichletC(_BCBase):
    def __init__(self, N, dealias=None):
        GalerkinChebyshev.__init__(self, N, dealias)
        self.id = "DC"
        self.is_bc = True

    def _stencil(self):
        S = np.zeros((self.N, 2))
        S[0, 0], S[1, 0] = 0.5, -0.5
        S[0, 1], S[1, 1] = 0.5, 0.5
        return S


class NeumannC(_BCBase):
    def __init__(self, N, dealias=None):
        GalerkinChebyshev.__init__(self, N, dealias)
        self.id = "NC"
        self.is_bc = True

    def _stencil(self):
        S = np.zeros((self.N, 2))
        S[0, 0], S[1, 0] = 0.5, -1 / 8
        S[0, 1], S[1, 1] = 0.5, 1 / 8
        return S


_KEYS = {
    "CH": Chebyshev, "Chebyshev": Chebyshev,
    "CD": ChebDirichlet, "ChebDirichlet": ChebDirichlet,
    "CN": ChebNeumann, "ChebNeumann": ChebNeumann,
    "CDN": ChebDirichletNeumann, "ChebDirichletNeumann": ChebDirichletNeumann,
    "DC": DirichletC, "DirichletC": DirichletC,
    "NC": NeumannC, "NeumannC": NeumannC,
}


def Base(N, key, *args, **kwargs):
    """Basis factory by key (pypde/bases/spectralbase.py:12-45)."""
    if key not in _KEYS:
        raise ValueError("Key {:} not available.".format(key))
    return _KEYS[key](N, *args, **kwargs)
