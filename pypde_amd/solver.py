"""Linear-solver plans on the GPU (drop-in for pypde/solver/plans.py,
solverplan.py, integrator.py, utils.py and linalg/pdma.py).

A plan is built once on the host in float64 (matrices, LU sweeps, per-line
Poisson factor tables) and applied on the device:
  PlanRHS  banded operators (|offset| <= 4)  -> pypde_banded_apply (csr order)
           other (dense) operators            -> pypde_dgemm
  fdma                                        -> pypde_fdma_solve
  twodma                                      -> pypde_twodma_solve
  poisson                                     -> pypde_poisson_solve
  numpy (general dense)                       -> inverse precomputed, pypde_dgemm
"""
import numpy as np
import scipy.sparse as sp
import torch

from . import ops


# ---------------------------------------------------------------------------
# normalised recurrence tables (reciprocal pivots, zero-padded chain ends)
# ---------------------------------------------------------------------------
def rec_from_fdma(l, d, u1, u2):
    """Four-diagonal LU tables (plans.py:226-234) -> (LM, R, MU, NU)."""
    n = d.shape[-1]
    lead = d.shape[:-1]
    LM = np.zeros(lead + (n,))
    MU = np.zeros(lead + (n,))
    NU = np.zeros(lead + (n,))
    LM[..., 2:] = -l[..., : n - 2]
    R = 1.0 / d
    MU[..., : n - 2] = -u1[..., : n - 2] / d[..., : n - 2]
    NU[..., : n - 4] = -u2[..., : n - 4] / d[..., : n - 4]
    return LM, R, MU, NU


def rec_from_poisson(tab, sing):
    """Per-line Poisson factors (poisson_line_tables) -> per-line normalised
    tables laid out [n][nlines]; singular lines act on positions 1.. only and
    return x_0 = 0 (fdma.f90:176-179)."""
    nl, _, n = tab.shape
    LM, R, MU, NU = (np.zeros((nl, n)) for _ in range(4))
    for j in range(nl):
        d, u1, u2, l = tab[j]
        if sing[j]:
            m = n - 1
            a, b, c, e = rec_from_fdma(l[: m - 2], d[:m], u1[: m - 2], u2[: m - 4])
            LM[j, 1:], R[j, 1:], MU[j, 1:], NU[j, 1:] = a, b, c, e
            R[j, 0] = 0.0
        else:
            a, b, c, e = rec_from_fdma(l[: n - 2], d[:n], u1[: n - 2], u2[: n - 4])
            LM[j], R[j], MU[j], NU[j] = a, b, c, e
    return tuple(np.ascontiguousarray(t.T) for t in (LM, R, MU, NU))


def _fdma_factor_diagonals(A):
    """Diagonals -2, 0, 2, 4 of a dense four-diagonal matrix, LU-factored
    (plans.py:226-234)."""
    l, d, u1, u2 = (np.diag(A, k).copy() for k in (-2, 0, 2, 4))
    fdma_lu(l, d, u1, u2)
    return l, d, u1, u2


def rec_reversed_transpose(A):
    """Normalised tables of x = A^-T b solved on the reversed line.  With J
    the reversal, J A^T J has A's band offsets -2..4, so (J A^T J)(J x) = J b is
    a four-diagonal solve of the usual form over the line read backwards."""
    A = _dense(A)
    return rec_from_fdma(*_fdma_factor_diagonals(A[::-1, ::-1].T))


def rec_poisson_reversed_transpose(A, C, alpha, singular):
    """Per-line tables of x_j = (A + alpha_j C)^-T b_j on the reversed line,
    the transpose of the line solves of poisson_line_tables.  A singular line
    (|alpha_j| < 1e-10 with singular set) solves the trailing n-1 block and sets
    x_0 = 0; its transpose maps b_0 to nothing and returns 0 there, so on the
    reversed line it solves the leading n-1 block and returns 0 at position
    n-1 (R = 0).  Laid out [n][nlines] per table."""
    A, C = _dense(A), _dense(C)
    Ar, Cr = A[::-1, ::-1].T, C[::-1, ::-1].T
    alpha = np.asarray(alpha, dtype=float)
    n, nl = A.shape[0], alpha.shape[0]
    LM, R, MU, NU = (np.zeros((nl, n)) for _ in range(4))
    for j in range(nl):
        nn = n - 1 if (singular and abs(alpha[j]) < 1e-10) else n
        M = Ar[:nn, :nn] + alpha[j] * Cr[:nn, :nn]
        LM[j, :nn], R[j, :nn], MU[j, :nn], NU[j, :nn] = rec_from_fdma(*_fdma_factor_diagonals(M))
    return tuple(np.ascontiguousarray(t.T) for t in (LM, R, MU, NU))


def rec_from_tdma(s, a, b, w, den):
    """Galerkin projection (S^T S)^-1 S^T (tdma.f90:56-107) -> (pre, F, LM, MU)."""
    M = b.shape[0]
    pre = np.array(s, dtype=float)
    F = np.empty(M)
    LM = np.zeros(M)
    F[:2] = 1.0 / b[:2]
    F[2:] = 1.0 / den[2:]
    LM[2:] = -a[: M - 2] / den[2:]
    MU = np.zeros(M)
    MU[: M - 2] = -w[: M - 2]
    return pre, F, LM, MU


def rec_from_twodma(d, u):
    n = d.shape[0]
    R = 1.0 / d
    MU = np.zeros(n)
    MU[: n - 2] = -u[: n - 2] / d[: n - 2]
    return R, MU


class Flags:
    def __init__(self, **kwargs):
        self.flags = dict(**kwargs)


def _dense(A):
    if sp.issparse(A):
        return np.asarray(A.toarray(), dtype=float)
    return np.asarray(A, dtype=float)


def banded_diagonals(A, max_off=4):
    """(offsets, vals) if every nonzero of A lies within |col-row| <= max_off."""
    A = _dense(A)
    rows, cols = np.nonzero(A)
    offs = sorted(set((cols - rows).tolist()))
    if not offs or min(offs) < -max_off or max(offs) > max_off:
        return None
    m = A.shape[0]
    vals = np.zeros((len(offs), m))
    for q, o in enumerate(offs):
        for r in range(m):
            c = r + o
            if 0 <= c < A.shape[1]:
                vals[q, r] = A[r, c]
    return offs, vals


class MetaPlan:
    def __init__(self, A, ndim, axis, **kwargs):
        self.A = A
        self.ndim = ndim
        self.axis = axis
        self.N = A.shape[1]
        self._flags = Flags()
        self.flags.update({"axis": axis})

    @property
    def flags(self):
        return self._flags.flags

    def _check_b(self, b):
        """Shape checks of plans.py:40-51; a leading batch dimension is allowed
        (axis then counts within the trailing ndim dimensions)."""
        assert isinstance(b, torch.Tensor)
        assert b.dim() in (self.ndim, self.ndim + 1), (
            "Dimensionality mismatch: ndim {:4d} | b.ndim {:4d}. Check ndim.".format(self.ndim, b.dim()))
        n = b.shape[b.dim() - self.ndim + self.axis]
        assert self.N == n, "Shape mismatch. : N {:4d} | b.shape[axis] {:4d}. Check axis.".format(
            self.N, n)


class PlanRHS(MetaPlan):
    """b -> A b along axis (plans.py:54-74)."""

    def __init__(self, A, ndim, axis):
        MetaPlan.__init__(self, A=A, ndim=ndim, axis=axis)
        self.flags.update({"method": "multiply"})
        if np.isscalar(A):
            self.scalar = float(A)
            return
        self.scalar = None
        band = banded_diagonals(A)
        self.M_out = _dense(A).shape[0]
        if band is not None:
            self.offsets, vals = band
            self.vals = ops.to_dev(vals)
            self.kind = "banded"
        else:
            self.mat = ops.to_dev(_dense(A))
            self.kind = "dense"

    def solve(self, b):
        b = ops.to_dev(b)
        if self.scalar is not None:
            return self.scalar * b
        self._check_b(b)
        if self.kind == "banded":
            return ops.banded_apply(self.offsets, self.vals, b, self.axis, self.M_out)
        return ops.matmul_axis(self.mat, b, self.axis)


def fdma_lu(l, d, u1, u2):
    """Plan-time sweep of the four-diagonal LU (plans.py:226-234), in place."""
    n = d.shape[0]
    for i in range(2, n):
        l[..., i - 2] = l[..., i - 2] / d[..., i - 2]
        d[..., i] = d[..., i] - l[..., i - 2] * u1[..., i - 2]
        if i < n - 2:
            u1[..., i] = u1[..., i] - l[..., i - 2] * u2[..., i - 2]


class Plan_fdma(MetaPlan):
    """A x = b, A with diagonals -2, 0, 2, 4 (plans.py:184-224)."""

    def __init__(self, A, ndim, axis):
        MetaPlan.__init__(self, A=A, ndim=ndim, axis=axis)
        self.flags.update({"method": "fdma"})
        A = _dense(A)
        l = np.diag(A, -2).copy()
        d = np.diag(A, 0).copy()
        u1 = np.diag(A, 2).copy()
        u2 = np.diag(A, 4).copy()
        fdma_lu(l, d, u1, u2)
        self.l, self.d, self.u1, self.u2 = l, d, u1, u2
        self.tabs = tuple(ops.to_dev(t) for t in (l, d, u1, u2))
        LM, R, MU, NU = rec_from_fdma(l, d, u1, u2)
        self.rec = ops.RecTables(LM=LM, R=R, MU=MU, NU=NU)

    def solve(self, b, exact=False):
        """In-place solve; exact=True runs the reference's sweep statement by
        statement (bit-identical), the default the register-resident kernel."""
        b = ops.to_dev(b)
        self._check_b(b)
        if self.ndim > 2:
            raise NotImplementedError("Plan_fdma supports only ndim<3.")
        if exact:
            return ops.fdma_solve(self.tabs, b, self.axis)
        return ops.recurrence_solve(self.rec, b, self.axis)


class Plan_twodma(MetaPlan):
    """A x = b, A with diagonals 0, 2 (plans.py:137-181)."""

    def __init__(self, A, ndim, axis):
        MetaPlan.__init__(self, A=A, ndim=ndim, axis=axis)
        self.flags.update({"method": "twodma"})
        A = _dense(A)
        d, u = np.diag(A, 0).copy(), np.diag(A, 2).copy()
        self.d = ops.to_dev(d)
        self.u = ops.to_dev(u)
        R, MU = rec_from_twodma(d, u)
        self.rec = ops.RecTables(R=R, MU=MU)

    def solve(self, b, exact=False):
        b = ops.to_dev(b)
        self._check_b(b)
        if self.ndim > 2:
            raise NotImplementedError("Plan_twodma supports only ndim<3.")
        if exact or b.shape[b.dim() - self.ndim + self.axis] < 3:
            return ops.twodma_solve(self.d, self.u, b, self.axis)
        return ops.recurrence_solve(self.rec, b, self.axis)


class Plan_numpy(MetaPlan):
    """General dense A x = b (plans.py:122-134); A^-1 is formed once."""

    def __init__(self, A, ndim, axis):
        MetaPlan.__init__(self, A=A, ndim=ndim, axis=axis)
        self.flags.update({"method": "numpy"})
        self.inv = ops.to_dev(np.linalg.inv(_dense(A)))

    def solve(self, b):
        b = ops.to_dev(b)
        self._check_b(b)
        return ops.matmul_axis(self.inv, b, self.axis)


def poisson_line_tables(A, C, alpha, singular):
    """Per-line factor tables of (A + alpha_i C) exactly as fdma.f90:173-184
    forms and factors them on every call (init_fdma :102-143), computed once.
    Returns tab [nlines, 4, n] (d | u1 | u2 | l) and the singular-line flags."""
    A, C = _dense(A), _dense(C)
    alpha = np.asarray(alpha, dtype=float)
    n = A.shape[0]
    nl = alpha.shape[0]
    sing = np.zeros(nl, dtype=np.int32)
    if singular:
        sing[np.abs(alpha) < 1e-10] = 1
    tab = np.zeros((nl, 4, n))
    for reduced in (0, 1):
        idx = np.nonzero(sing == reduced)[0]
        if idx.size == 0:
            continue
        o = reduced
        nn = n - o
        Ar, Cr = A[o:, o:], C[o:, o:]
        lam = alpha[idx][:, None]
        # (A + C*lam) elementwise, then the diagonals init_fdma extracts
        d = np.diag(Ar)[None, :] + np.diag(Cr)[None, :] * lam
        u1 = np.diag(Ar, 2)[None, :] + np.diag(Cr, 2)[None, :] * lam
        u2 = np.diag(Ar, 4)[None, :] + np.diag(Cr, 4)[None, :] * lam
        l = np.diag(Ar, -2)[None, :] + np.diag(Cr, -2)[None, :] * lam
        d, u1, u2, l = (np.ascontiguousarray(t) for t in (d, u1, u2, l))
        for i in range(2, nn):
            l[:, i - 2] = l[:, i - 2] / d[:, i - 2]
            d[:, i] = d[:, i] - l[:, i - 2] * u1[:, i - 2]
            if i < nn - 2:
                u1[:, i] = u1[:, i] - l[:, i - 2] * u2[:, i - 2]
        tab[idx, 0, :nn] = d
        tab[idx, 1, :nn - 2] = u1
        tab[idx, 2, :nn - 4] = u2
        tab[idx, 3, :nn - 2] = l
    return tab, sing


class Plan_Poisson(MetaPlan):
    """(A + alpha_i C) x_i = b_i per line (plans.py:254-341)."""

    def __init__(self, A, alpha, C, ndim, axis, singular=False):
        assert ndim == 2
        MetaPlan.__init__(self, A=A, ndim=ndim, axis=axis)
        self.flags.update({"method": "poisson"})
        self.alpha = np.asarray(alpha)
        self.C = C
        self.singular = singular
        # the reference's singular branch exists for axis 0 only (fdma.f90:176)
        tab, sing = poisson_line_tables(A, C, self.alpha, singular and axis == 0)
        self.tab = ops.to_dev(tab)
        self.sing = torch.as_tensor(sing, device=ops.device())
        LM, R, MU, NU = rec_from_poisson(tab, sing)
        self.rec = ops.RecTables(LM=LM, R=R, MU=MU, NU=NU, per_line=True)

    def solve(self, b, exact=False):
        b = ops.to_dev(b)
        other = b.shape[b.dim() - 1 - self.axis]
        if self.alpha.size != other:
            raise ValueError("Size of eigenvalue {:3} array does not match to b {:3}!".format(
                self.alpha.size, other))
        if exact:
            return ops.poisson_solve(self.tab, self.sing, b, self.axis)
        return ops.recurrence_solve(self.rec, b, self.axis)


def PlanLHS(A, ndim, axis, method, **kwargs):
    """Plan factory (plans.py:77-119)."""
    all_method = {"numpy": Plan_numpy, "twodma": Plan_twodma, "fdma": Plan_fdma,
                  "poisson": Plan_Poisson, "multiply": PlanRHS}
    if method not in all_method:
        raise ValueError("Method name {:} not found in: {:}".format(method, all_method.keys()))
    return all_method[method](A=A, ndim=ndim, axis=axis, **kwargs)


class SolverPlan:
    """Ordered rhs / old / lhs plan chains (solverplan.py:4-100)."""

    def __init__(self):
        self.plan_for_lhs = []
        self.plan_for_rhs = []
        self.plan_for_old = []

    def add_rhs(self, plan):
        assert isinstance(plan, PlanRHS)
        self.plan_for_rhs.append(plan)

    def add_old(self, plan):
        assert isinstance(plan, PlanRHS)
        self.plan_for_old.append(plan)

    def add_lhs(self, plan):
        assert isinstance(plan, MetaPlan)
        self.plan_for_lhs.append(plan)

    def solve_rhs(self, b):
        for plan in self.plan_for_rhs:
            b = plan.solve(b)
        return b

    def solve_old(self, b):
        for plan in self.plan_for_old:
            b = plan.solve(b)
        return b

    def solve_lhs(self, b):
        for plan in self.plan_for_lhs:
            b = plan.solve(b)
        return b

    def show_plan(self):
        print("Plans RHS:")
        for i, p in enumerate(self.plan_for_rhs):
            print(i + 1, ")", self._str_plan(p))
        print("")
        if self.plan_for_old:
            print("Plans RHS (#2):")
            for i, p in enumerate(self.plan_for_old):
                print(i + 1, ")", self._str_plan(p))
            print("")
        print("Plans LHS:")
        for i, p in enumerate(self.plan_for_lhs):
            print(i + 1, ")", self._str_plan(p))
        print("")

    @staticmethod
    def _str_plan(plan):
        return "Apply method '{:s}' along axis {:1d} ".format(plan.flags["method"], plan.flags["axis"])


class PDMA:
    """Pentadiagonal (-4,-2,0,2,4) solver (pypde/solver/linalg/pdma.py:4-93):
    PDMA_LU on the host, the sweeps on the device (axis 0 of the input)."""

    def __init__(self, mat, neumann=False):
        self.mat = _dense(mat)
        self.neumann = neumann

    def init(self):
        B = self.mat
        n = B.shape[1]
        d0 = np.diag(B, 0).copy()
        d1 = np.diag(B, 2).copy()
        d2 = np.diag(B, 4).copy()
        if self.neumann:
            d0[0], d1[0], d2[0] = 1, 0, 0
        l1 = np.diag(B, -2).copy()
        l2 = np.diag(B, -4).copy()
        self.PDMA_LU(l2, l1, d0, d1, d2)
        self.tabs = tuple(ops.to_dev(t) for t in (l2, l1, d0, d1, d2))
        self.N = n

    @staticmethod
    def PDMA_LU(a, b, d, e, f):
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
        for i in (n - 4, n - 3):
            lam = b[i] / d[i]
            d[i + k] -= lam * e[i]
            b[i] = lam

    def solve(self, x, axis=0):
        return ops.pdma_solve(self.tabs, x, axis)


def eigdecomp(A):
    """Eigen-decomposition sorted by descending eigenvalue (solver/utils.py:13-29)."""
    w, Q = np.linalg.eig(A)
    argsort = np.argsort(w)[::-1]
    w = w[argsort]
    Q = Q[:, argsort]
    Qi = np.linalg.inv(Q)
    return w, Q, Qi


class Integrator:
    """Time loop (pypde/solver/integrator.py:4-44)."""

    def __init__(self):
        self.time = 0.0

    def update(self):
        raise NotImplementedError

    def callback(self):
        pass

    def iterate(self, maxtime, callback=True, *args, **kwargs):
        eps = 1e-3 * self.dt
        while (self.time + eps) < maxtime:
            self.update(*args, **kwargs)
            self.update_time()
            if self.tsave is not None:
                if (self.time + eps) % self.tsave < self.dt * 0.5:
                    self.save()
                    print("Time: {:5.3f}".format(self.time))
                    if callback:
                        self.callback()

    def update_time(self):
        from .field import Field, MultiField
        if isinstance(self.field, Field):
            self.field.t += self.dt
        elif isinstance(self.field, MultiField):
            self.field.update_time(self.dt)
        else:
            raise ValueError("fields must be of type Field or MultiField.")
        self.time += self.dt

    def save(self):
        self.field.save()
