"""Helmholtz / Poisson plan templates (drop-in for pypde/templates/hholtz.py and
pypde/templates/poisson.py)."""
import numpy as np

from .field import Field
from .solver import PlanLHS, PlanRHS, SolverPlan, eigdecomp


def _mats(basis):
    S = basis.S_sp
    B = basis.family.B(2, 2)
    I = basis.family.I(2)
    return S, B, I


def solverplan_hholtz1d(bases, lam):
    """(B - lam I) S v = B rhs + B S vold (hholtz.py:4-39)."""
    field = Field(bases)
    assert field.ndim == 1
    Sx, Bx, Ix = _mats(field.xs[0])
    Ax = Bx @ Sx - lam * Ix @ Sx
    solver = SolverPlan()
    solver.add_rhs(PlanRHS(Bx, ndim=1, axis=0))
    solver.add_old(PlanRHS(Bx @ Sx, ndim=1, axis=0))
    solver.add_lhs(PlanLHS(Ax, ndim=1, axis=0, method="fdma"))
    return solver


def hholtz_matrices(basis, lam, scale):
    Sx, Bx, Ix = _mats(basis)
    Ax = Bx @ Sx - lam * (1.0 / scale ** 2.0) * Ix @ Sx
    return Bx, np.asarray(Bx @ Sx), np.asarray(Ax)


def solverplan_hholtz2d_adi(bases, lam, scale=(1, 1)):
    """ADI Helmholtz plan (hholtz.py:42-93): Ax^-1 (Bx R By^T + BxSx V By Sy^T) Ay^-T."""
    field = Field(bases)
    assert field.ndim == 2
    Bx, BSx, Ax = hholtz_matrices(field.xs[0], lam, scale[0])
    By, BSy, Ay = hholtz_matrices(field.xs[1], lam, scale[1])
    solver = SolverPlan()
    solver.add_rhs(PlanRHS(Bx, ndim=2, axis=0))
    solver.add_old(PlanRHS(BSx, ndim=2, axis=0))
    solver.add_lhs(PlanLHS(Ax, ndim=2, axis=0, method="fdma"))
    solver.add_rhs(PlanRHS(By, ndim=2, axis=1))
    solver.add_old(PlanRHS(BSy, ndim=2, axis=1))
    solver.add_lhs(PlanLHS(Ay, ndim=2, axis=1, method="fdma"))
    return solver


def solverplan_poisson1d(bases, singular=False):
    """I S v = B rhs (poisson.py:4-40)."""
    field = Field(bases)
    assert field.ndim == 1
    Sx, Bx, Ix = _mats(field.xs[0])
    Ax = np.asarray(Ix @ Sx)
    if singular:
        assert Ax[0, 0] == 0, "Matrix does not look singular"
        Ax[0, 0] += 1e-20
    solver = SolverPlan()
    solver.add_rhs(PlanRHS(Bx, ndim=1, axis=0))
    solver.add_lhs(PlanLHS(Ax, ndim=1, axis=0, method="twodma"))
    return solver


def poisson_matrices(bases, singular=False, scale=(1, 1)):
    """Matrices of the eigen-split Poisson plan (poisson.py:43-110)."""
    Sx, Bx, Ix = _mats(bases[0])
    Ax = np.asarray(Ix @ Sx * (1.0 / scale[0] ** 2.0))
    Cx = np.asarray(Bx @ Sx)
    Sy, By, Iy = _mats(bases[1])
    Ay = np.asarray(Iy @ Sy * (1.0 / scale[1] ** 2.0))
    Cy = np.asarray(By @ Sy)
    CyI = np.linalg.inv(Cy)
    wy, Qy, QyI = eigdecomp(CyI @ Ay)
    if singular:
        wy[0] += 1e-20
    Hy = QyI @ CyI @ By
    return dict(Ax=Ax, Cx=Cx, Ay=Ay, Cy=Cy, Bx=Bx, By=By, wy=np.real(wy), Qy=np.real(Qy),
                QyI=np.real(QyI), Hy=np.real(Hy))


def solverplan_poisson2d(bases, singular=False, scale=(1, 1)):
    """Eigen-split Poisson plan; the LHS is always built with singular=True
    like the reference (poisson.py:106)."""
    field = Field(bases)
    assert field.ndim == 2
    m = poisson_matrices(field.xs, singular, scale)
    solver = SolverPlan()
    solver.add_rhs(PlanRHS(m["Bx"], ndim=2, axis=0))
    solver.add_rhs(PlanRHS(m["Hy"], ndim=2, axis=1))
    solver.add_lhs(PlanLHS(m["Ax"], alpha=m["wy"], C=m["Cx"], ndim=2, axis=0, method="poisson",
                           singular=True))
    solver.add_lhs(PlanLHS(m["Qy"], ndim=2, axis=1, method="multiply"))
    solver.mats = m
    return solver
