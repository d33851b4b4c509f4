"""Rayleigh-Benard convection in a 2-D box on the GPU (drop-in for
navier/rbc2d.py:28-434).

Same constructor (`NavierStokes(case, shape=, ra=, pr=, dt=, tsave=, dealias=,
integrator=, beta=, aspect=)`), fields (T, U, V, P, pres with .vhat / .v),
`update()`, `iterate()`, `eval_Nu()`.  The step runs on the fused engine
(pypde_amd.engine.RBCBatch, 19 launches per stage, optionally replayed from a
HIP graph) for both integrators and any beta; only RK3 or beta != 1 with an odd
dealiased y length falls back to the reference's operation sequence on the
drop-in GPU operators.
"""
import numpy as np
import torch

from .. import ops
from ..bases import Base
from ..engine import RBCBatch, transfer_function
from ..field import Field, FieldBC, MultiField
from ..field_operations import cheby_to_galerkin, convective_term, galerkin_to_cheby
from ..solver import Integrator
from ..templates import solverplan_hholtz2d_adi, solverplan_poisson2d
from .rbc2d_base import NavierStokesBase


class NavierStokes(NavierStokesBase, Integrator):
    """rbc: adiabatic sidewalls; linear: isothermal sidewalls with a linear
    profile; zero: isothermal sidewalls at zero temperature."""

    avail_cases = ["rbc", "linear", "zero"]

    def __init__(self, case="rbc", use_graph=True, exact_solves=False, **kwargs):
        if case not in self.avail_cases:
            raise ValueError("Specified case is not available: ", self.avail_cases)
        self.case = case
        # line solves as the kernels that keep the reference's LU statement
        # order (slower than the default register recurrences)
        self.exact_solves = exact_solves
        NavierStokesBase.__init__(self, **kwargs)
        Integrator.__init__(self)
        side = "CN" if case == "rbc" else "CD"
        N0, N1 = self.shape
        self.T = Field([Base(N0, side, dealias=3 / 2), Base(N1, "CD", dealias=3 / 2)])
        self.U = Field([Base(N0, "CD", dealias=3 / 2), Base(N1, "CD", dealias=3 / 2)])
        self.V = Field([Base(N0, "CD", dealias=3 / 2), Base(N1, "CD", dealias=3 / 2)])
        self.P = Field([Base(N0, "CN"), Base(N1, "CN")])
        self.pres = Field([Base(N0, "CH"), Base(N1, "CH")])
        self.field = MultiField([self.T, self.U, self.V, self.pres], ["temp", "ux", "uy", "pres"])
        self.use_graph = use_graph
        self.setup_solver()

    # -----------------------------------------------------------------------
    @property
    def fused(self):
        """True when the step runs on the fused engine."""
        return getattr(self, "engine", None) is not None

    def setup_solver(self):
        if self.integrator == "rk3":
            self.set_timestep_coefficients_rk3()
        else:
            self.set_timestep_coefficients_euler()
        keep = None
        if hasattr(self, "engine") or hasattr(self, "solver_U"):
            keep = self._get_state()
        try:
            engine = RBCBatch(self.shape, [dict(ra=self.ra, pr=self.pr, dt=self.dt, nu=self.nu,
                                                kappa=self.kappa)],
                              case=self.case, dealias=self.dealias, aspect=self.aspect,
                              beta=self.beta, integrator=self.integrator,
                              exact_solves=self.exact_solves)
        except NotImplementedError:
            # RK3 / beta != 1 with odd dealiased lengths: the op-by-op path
            engine = None
        if engine is not None:
            self.engine = engine
            e = self.engine
            self.Tbc, self.dTbcdz1, self.dTbcdz2, self.Tbc_cheby = (e.Tbc, e.dTbcdz1, e.dTbcdz2,
                                                                   e.Tbc_cheby)
            self._bind_views()
            self._graph_ready = False
        else:
            self.engine = None
            self.solver_U, self.solver_V, self.solver_T = [], [], []
            for rk in range(self.nstage):
                lu = self.dt * self.a[rk] * self.beta * self.nu
                lt = self.dt * self.a[rk] * self.beta * self.kappa
                self.solver_U.append(solverplan_hholtz2d_adi(self.U.xs, lu, self.scale))
                self.solver_V.append(solverplan_hholtz2d_adi(self.V.xs, lu, self.scale))
                self.solver_T.append(solverplan_hholtz2d_adi(self.T.xs, lt, self.scale))
            self.solver_P = solverplan_poisson2d(self.P.xs, singular=True, scale=self.scale)
            self.set_fieldbc()
        if keep is not None:
            self.set_state(**keep)

    def set_fieldbc(self):
        N0, N1 = self.shape
        if self.case == "zero":
            bc = np.zeros((2, N1))
            bc[0, :] = transfer_function(0.5, 0, -0.5, self.y, k=0.02)
            bc[1, :] = bc[0, :]
            self.Tbc = FieldBC(self.T.xs, axis=0)
        else:
            bc = np.zeros((N0, 2))
            bc[:, 0], bc[:, 1] = 0.5, -0.5
            self.Tbc = FieldBC(self.T.xs, axis=1)
        self.Tbc.add_bc(bc)
        self.dTbcdz2 = self.grad(self.Tbc, deriv=(0, 2))
        vhat = self.grad(self.Tbc, deriv=(0, 1))
        if self.dealias:
            self.dTbcdz1 = self.deriv_field.dealias.backward(vhat)
        else:
            self.dTbcdz1 = self.deriv_field.backward(vhat)
        self.Tbc_cheby = galerkin_to_cheby(self.Tbc.vhat, self.Tbc)

    def reset(self, reset_time=True):
        self.set_nu_kappa()
        self.setup_solver()
        if reset_time:
            self.reset_time()

    def reset_time(self):
        self.time = 0.0
        for field in self.field.fields:
            field.t = 0.0

    # -- engine state binding ------------------------------------------------
    def _bind_views(self):
        e = self.engine
        self.U.vhat, self.V.vhat, self.T.vhat = e.X[0, 0], e.X[0, 1], e.X[0, 2]
        self.pres.vhat = e.pres[0]
        self.P.vhat = e.P[0]
        self._bound = [t.data_ptr() for t in (self.U.vhat, self.V.vhat, self.T.vhat,
                                             self.pres.vhat, self.P.vhat)]

    def _sync_in(self):
        """Copy state tensors the user replaced back into the engine buffers."""
        e = self.engine
        dst = (e.X[0, 0], e.X[0, 1], e.X[0, 2], e.pres[0], e.P[0])
        src = (self.U.vhat, self.V.vhat, self.T.vhat, self.pres.vhat, self.P.vhat)
        changed = False
        for d, s, ptr in zip(dst, src, self._bound):
            if s.data_ptr() != ptr:
                d.copy_(ops.to_dev(s).reshape(d.shape))
                changed = True
        if changed:
            self._bind_views()

    def _get_state(self):
        return dict(T=self.T.vhat.clone(), U=self.U.vhat.clone(), V=self.V.vhat.clone(),
                    P=self.P.vhat.clone(), pres=self.pres.vhat.clone())

    def set_state(self, T=None, U=None, V=None, P=None, pres=None):
        for name, val in (("T", T), ("U", U), ("V", V), ("P", P), ("pres", pres)):
            if val is not None:
                f = getattr(self, name)
                if self.engine is not None:
                    f.vhat.copy_(ops.to_dev(val).reshape(f.vhat.shape))
                else:
                    f.vhat = ops.to_dev(val).reshape(f.shape_spectral).clone()

    # -- initial conditions (rbc2d.py:122-133) -------------------------------
    def set_temperature(self, amplitude=0.5, m=1):
        v = amplitude * np.sin(m * np.pi * self.xx) * np.cos(np.pi * self.yy)
        self.set_state(T=self.T.forward(v))

    def set_velocity(self, amplitude=0.5, m=1, n=1):
        x = (self.x - self.x[0]) / (self.x[-1] - self.x[0])
        y = (self.y - self.y[0]) / (self.y[-1] - self.y[0])
        xx, yy = np.meshgrid(x, y, indexing="ij")
        u = -amplitude * np.sin(m * np.pi * xx) * np.cos(n * np.pi * yy)
        v = amplitude * np.cos(m * np.pi * xx) * np.sin(n * np.pi * yy)
        self.set_state(U=self.U.forward(u), V=self.V.forward(v))

    # -- the time step --------------------------------------------------------
    def update(self):
        if self.engine is not None:
            self._sync_in()
            if self.use_graph and not self._graph_ready:
                self.engine.capture(1)
                self._graph_ready = True
            self.engine.step(1)
            return
        self._update_ops()

    # steps per captured graph for update_many: one graph launch costs about
    # 5 us of the 150 us 256x256 step, so blocks of 10 steps take 3% off
    GRAPH_BLOCK = 10

    def prepare_graphs(self):
        """Capture the one-step graph and the GRAPH_BLOCK-step graph now (the
        state is left unchanged), so that no later update() or update_many()
        call captures: a caller that times steps calls this first."""
        if self.engine is None or not self.use_graph:
            return
        self._sync_in()
        if not self._graph_ready:
            self.engine.capture(1)
            self._graph_ready = True
        if self.GRAPH_BLOCK not in self.engine.graphs:
            self.engine.capture(self.GRAPH_BLOCK)

    def update_many(self, n):
        """n steps at once, no host work between: replays of a graph of
        GRAPH_BLOCK steps, then single-step replays for the rest."""
        if self.engine is None:
            for _ in range(n):
                self.update()
            return
        self._sync_in()
        if self.use_graph:
            if not self._graph_ready:
                self.engine.capture(1)
                self._graph_ready = True
            if n >= self.GRAPH_BLOCK and self.GRAPH_BLOCK not in self.engine.graphs:
                self.engine.capture(self.GRAPH_BLOCK)
        self.engine.step(n)

    # op-by-op path: the reference's sequence (rbc2d.py:213-434) on GPU operators
    def conv_term(self, field, ux, uz, add_bc=None):
        return convective_term(field, ux, uz, deriv_field=self.deriv_field, add_bc=add_bc,
                               dealias=self.dealias, scale=self.scale)

    def _diffusion(self, rhs, field, coef, stage):
        if self.beta != 1.0:
            f = self.dt * self.a[stage] * (1 - self.beta) * coef
            rhs = rhs + f * self.grad(field, deriv=(2, 0))
            rhs = rhs + f * self.grad(field, deriv=(0, 2))
        return rhs

    def _update_ops(self):
        self.ux_old, self.uz_old = 0, 0
        for rk in range(self.nstage):
            That = galerkin_to_cheby(self.T.vhat, self.T) + self.Tbc_cheby
            if self.dealias:
                self.ux = self.U.dealias.backward(self.U.vhat)
                self.uz = self.V.dealias.backward(self.V.vhat)
            else:
                self.ux = self.U.backward(self.U.vhat)
                self.uz = self.V.backward(self.V.vhat)
            dt, a, b, c = self.dt, self.a[rk], self.b[rk], self.c[rk]
            # U
            rhs = -dt * a * self.grad(self.pres, deriv=(1, 0))
            rhs = rhs - dt * b * self.conv_term(self.U, self.ux, self.uz)
            if c != 0:
                rhs = rhs - dt * c * self.conv_term(self.U, self.ux_old, self.uz_old)
            rhs = self._diffusion(rhs, self.U, self.nu, rk)
            s = self.solver_U[rk]
            rhs = s.solve_rhs(rhs) + s.solve_old(self.U.vhat)
            self.U.vhat = s.solve_lhs(rhs)
            # V
            rhs = -dt * a * self.grad(self.pres, deriv=(0, 1))
            rhs = rhs - dt * b * self.conv_term(self.V, self.ux, self.uz)
            if c != 0:
                rhs = rhs - dt * c * self.conv_term(self.V, self.ux_old, self.uz_old)
            rhs = rhs + dt * a * That
            rhs = self._diffusion(rhs, self.V, self.nu, rk)
            s = self.solver_V[rk]
            rhs = s.solve_rhs(rhs) + s.solve_old(self.V.vhat)
            self.V.vhat = s.solve_lhs(rhs)
            # divergence, pressure, correction
            div = self.grad(self.U, deriv=(1, 0)) + self.grad(self.V, deriv=(0, 1))
            rhs = self.solver_P.solve_rhs(div)
            P = self.solver_P.solve_lhs(rhs)
            P[0, 0].zero_()  # a device fill (capturable; no host copy)
            self.P.vhat = P
            self.pres.vhat = self.pres.vhat - 1.0 * self.nu * div * self.beta
            self.pres.vhat = self.pres.vhat + 1.0 / (dt * a) * galerkin_to_cheby(self.P.vhat, self.P)
            dpdx = self.grad(self.P, deriv=(1, 0))
            dpdz = self.grad(self.P, deriv=(0, 1))
            self.U.vhat = self.U.vhat - cheby_to_galerkin(dpdx * 1.0, self.U)
            self.V.vhat = self.V.vhat - cheby_to_galerkin(dpdz * 1.0, self.V)
            # T
            rhs = -dt * b * self.conv_term(self.T, self.ux, self.uz, add_bc=self.uz * self.dTbcdz1)
            if c != 0:
                rhs = rhs - dt * c * self.conv_term(self.T, self.ux_old, self.uz_old,
                                                    add_bc=self.uz_old * self.dTbcdz1)
            rhs = rhs + dt * a * self.kappa * self.dTbcdz2
            rhs = self._diffusion(rhs, self.T, self.kappa, rk)
            s = self.solver_T[rk]
            rhs = s.solve_rhs(rhs) + s.solve_old(self.T.vhat)
            self.T.vhat = s.solve_lhs(rhs)
            self.ux_old, self.uz_old = self.ux, self.uz

    def divergence_velocity(self, u, v):
        return self.grad(u, deriv=(1, 0)) + self.grad(v, deriv=(0, 1))
