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
        assert isinstance(b, torch.Tensor)
        assert self.ndim == b.dim(), "Dimensionality mismatch: ndim {:4d} | b.ndim {:4d}. Check ndim.".format(
            self.ndim, b.dim())
        assert self.N == b.shape[self.axis], "Shape mismatch. : N {:4d} | b.shape[axis] {:4d}. Check axis.".format(
            self.N, b.shape[self.axis])


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

    def solve(self, b):
        b = ops.to_dev(b)
        self._check_b(b)
        if self.ndim > 2:
            raise NotImplementedError("Plan_fdma supports only ndim<3.")
        return ops.fdma_solve(self.tabs, b, self.axis)


class Plan_twodma(MetaPlan):
    """A x = b, A with diagonals 0, 2 (plans.py:137-181)."""

    def __init__(self, A, ndim, axis):
        MetaPlan.__init__(self, A=A, ndim=ndim, axis=axis)
        self.flags.update({"method": "twodma"})
        A = _dense(A)
        self.d = ops.to_dev(np.diag(A, 0).copy())
        self.u = ops.to_dev(np.diag(A, 2).copy())

    def solve(self, b):
        b = ops.to_dev(b)
        self._check_b(b)
        if self.ndim > 2:
            raise NotImplementedError("Plan_twodma supports only ndim<3.")
        return ops.twodma_solve(self.d, self.u, b, self.axis)


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

    def solve(self, b):
        b = ops.to_dev(b)
        other = b.shape[1 - self.axis]
        if self.alpha.size != other:
            raise ValueError("Size of eigenvalue {:3} array does not match to b {:3}!".format(
                self.alpha.size, other))
        return ops.poisson_solve(self.tab, self.sing, b, self.axis)


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
