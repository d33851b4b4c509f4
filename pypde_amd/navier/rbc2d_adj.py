"""Adjoint-descent iterations towards steady states on the GPU (drop-in for
navier/rbc2d_adj.py:15-356).

Same constructor (`NavierStokesAdjoint(case, shape=, ra=, pr=, dt=, ...)`),
fields (T, U, V, P, pres; adjoint fields TA, UA, VA), inner solver `NS`,
`update()`, `iterate()`.  One stage:
  1. R = (step(X) - X)/dt from the inner NavierStokes (the fused engine; its
     pressure carries over between calls, as in the reference), written back
     into NS.T/U/V;
  2. adjoint fields by Poisson smoothing: nabla^2 UA = R_U / nu,
     nabla^2 VA = R_V / nu, nabla^2 TA = R_T / kappa;
  3. U, V move along -a grad(pres), the convective terms of UA / VA with the
     current velocities and the adjoint coupling terms, and a R;
  4. the velocity is projected (Poisson, pres update, correction);
  5. T moves along the convection of TA, a R_T and the buoyancy a VA.
Arrays are float64 device tensors.
"""
import numpy as np
import torch

from .. import ops
from ..bases import Base
from ..field import Field, MultiField
from ..field_operations import cheby_to_galerkin, conv_term, galerkin_to_cheby
from ..solver import Integrator
from ..templates import solverplan_poisson2d
from .rbc2d import NavierStokes
from .rbc2d_base import NavierStokesBase


class NavierStokesAdjoint(NavierStokesBase, Integrator):
    avail_cases = ["rbc", "linear", "zero"]

    def __init__(self, case="rbc", use_graph=True, **kwargs):
        if case not in self.avail_cases:
            raise ValueError("Specified case is not available: ", self.avail_cases)
        self.case = case
        # update() as a HIP graph: the first call runs op by op (and tunes every
        # kernel choice), the second captures one update, later calls replay it
        self.use_graph = use_graph
        self._graph = None
        self._calls = 0
        self._capturing = False
        Integrator.__init__(self)
        NavierStokesBase.__init__(self, **kwargs)
        side = "CN" if case == "rbc" else "CD"
        N0, N1 = self.shape
        # inner solver for the residual (rbc2d_adj.py:44)
        self.NS = NavierStokes(case=case, **self.CONFIG)
        # the captured update runs the inner step as the fused engine's launches
        self.use_graph = self.use_graph and self.NS.engine is not None

        def fld(kx, ky):
            return Field([Base(N0, kx, dealias=3 / 2), Base(N1, ky, dealias=3 / 2)])

        self.T, self.U, self.V = fld(side, "CD"), fld("CD", "CD"), fld("CD", "CD")
        self.P = Field([Base(N0, "CN"), Base(N1, "CN")])
        self.pres = Field([Base(N0, "CH"), Base(N1, "CH")])
        self.field = MultiField([self.T, self.U, self.V], ["temp", "ux", "uy"])
        self.TA, self.UA, self.VA = fld(side, "CD"), fld("CD", "CD"), fld("CD", "CD")
        self.setup_solver()
        self.Tbc = self.NS.Tbc
        self.Tbc.backward()
        # conductive profile on the physical (dealiased) grid (rbc2d_adj.py:112-117)
        self.deriv_field.vhat = galerkin_to_cheby(self.NS.Tbc.vhat, self.NS.Tbc)
        if self.dealias:
            self.temp_bc = self.deriv_field.dealias.backward(self.deriv_field.vhat)
        else:
            self.temp_bc = self.deriv_field.backward(self.deriv_field.vhat)

    def reset_time(self):
        self.time = 0.0
        for field in self.field.fields:
            field.time = 0.0

    def set_temperature(self, amplitude=0.5):
        v = amplitude * np.sin(0.5 * np.pi * self.xx) * np.cos(0.5 * np.pi * self.yy)
        self.T.vhat = self.T.forward(v)

    def setup_solver(self):
        """Time-step coefficients and the pressure plan of the inner solver;
        non-singular Poisson plans on each field's bases for the smoothing
        (rbc2d_adj.py:129-144)."""
        self.a, self.b, self.c = self.NS.a, self.NS.b, self.NS.c
        self.nstage = self.NS.nstage
        self.solver_P = solverplan_poisson2d(self.P.xs, singular=True, scale=self.scale)
        self.nabla_U = solverplan_poisson2d(self.U.xs, singular=False, scale=self.scale)
        self.nabla_V = solverplan_poisson2d(self.V.xs, singular=False, scale=self.scale)
        self.nabla_T = solverplan_poisson2d(self.T.xs, singular=False, scale=self.scale)

    def save(self):
        for f in (self.T, self.P, self.U, self.V):
            f.save()

    # -- terms ---------------------------------------------------------------
    def conv(self, field, u, deriv):
        return conv_term(field, u, deriv=deriv, deriv_field=self.deriv_field,
                         dealias=self.dealias, scale=self.scale)

    def _conv_adj(self, deriv, fieldx, fieldz, fieldT, ux, uz, temp):
        conv = self.conv(fieldx, ux, deriv)
        conv = conv + self.conv(fieldz, uz, deriv)
        conv = conv + self.conv(fieldT, temp, deriv)
        conv = conv + self.conv(fieldT, self.temp_bc, deriv)
        if self.dealias:
            return self.deriv_field.dealias.forward(conv)
        return self.deriv_field.forward(conv)

    def conv_term_adj_ux(self, fieldx, fieldz, fieldT, ux, uz, temp, add_bc=None):
        """x-derivative coupling terms (rbc2d_adj.py:162-174)."""
        return self._conv_adj((1, 0), fieldx, fieldz, fieldT, ux, uz, temp)

    def conv_term_adj_uz(self, fieldx, fieldz, fieldT, ux, uz, temp, add_bc=None):
        """y-derivative coupling terms (rbc2d_adj.py:176-187)."""
        return self._conv_adj((0, 1), fieldx, fieldz, fieldT, ux, uz, temp)

    def update_NS(self):
        """Residual of the inner step, then the smoothed adjoint fields
        (rbc2d_adj.py:189-205)."""
        NS = self.NS
        NS.set_state(U=self.U.vhat, V=self.V.vhat, T=self.T.vhat)
        if self._capturing and NS.engine is not None:
            NS._sync_in()
            NS.engine.step_eager()  # a graph replay cannot be captured
        else:
            NS.update()
        rU = (NS.U.vhat - self.U.vhat) / self.dt
        rV = (NS.V.vhat - self.V.vhat) / self.dt
        rT = (NS.T.vhat - self.T.vhat) / self.dt
        NS.set_state(U=rU, V=rV, T=rT)
        for plan, r, src, dst, coef in ((self.nabla_U, rU, NS.U, self.UA, self.nu),
                                        (self.nabla_V, rV, NS.V, self.VA, self.nu),
                                        (self.nabla_T, rT, NS.T, self.TA, self.kappa)):
            rhs = plan.solve_rhs(galerkin_to_cheby(r, src))
            dst.vhat = plan.solve_lhs(rhs) / coef

    def _velocity_update(self, stage, deriv, field, own_A, resid):
        a, b, c = self.a[stage], self.b[stage], self.c[stage]
        rhs = -a * self.grad(self.pres, deriv=deriv)
        rhs = rhs + b * self.NS.conv_term(own_A, self.ux, self.uz)
        rhs = rhs + b * self._conv_adj(deriv, self.UA, self.VA, self.TA, self.ux, self.uz,
                                       self.temp)
        if c != 0:
            rhs = rhs + c * self.NS.conv_term(own_A, self.ux_old, self.uz_old)
            rhs = rhs + c * self._conv_adj(deriv, self.UA, self.VA, self.TA, self.ux_old,
                                           self.uz_old, self.temp_old)
        rhs = rhs + a * galerkin_to_cheby(resid, field)
        field.vhat = field.vhat + self.dt * cheby_to_galerkin(rhs, field)

    def update_U(self, stage):
        """rbc2d_adj.py:207-232."""
        self._velocity_update(stage, (1, 0), self.U, self.UA, self.NS.U.vhat)

    def update_V(self, stage):
        """rbc2d_adj.py:234-257."""
        self._velocity_update(stage, (0, 1), self.V, self.VA, self.NS.V.vhat)

    def update_T(self, stage):
        """rbc2d_adj.py:259-281."""
        a, b, c = self.a[stage], self.b[stage], self.c[stage]
        rhs = b * self.NS.conv_term(self.TA, self.ux, self.uz)
        if c != 0:
            rhs = rhs + c * self.NS.conv_term(self.TA, self.ux_old, self.uz_old)
        rhs = rhs + a * galerkin_to_cheby(self.NS.T.vhat, self.T)
        rhs = rhs + a * galerkin_to_cheby(self.VA.vhat, self.VA)
        self.T.vhat = self.T.vhat + self.dt * cheby_to_galerkin(rhs, self.T)

    def update_P(self, div, singular=True):
        rhs = self.solver_P.solve_rhs(div)
        P = self.solver_P.solve_lhs(rhs)
        if singular:
            P = P.clone()
            P[0, 0].zero_()  # a device fill (capturable; no host copy)
        self.P.vhat = P

    def update_pres(self, div, stage):
        self.pres.vhat = self.pres.vhat + galerkin_to_cheby(self.P.vhat, self.P) / (
            self.dt * self.a[stage])

    def update_velocity(self, p, u, v, fac=1.0):
        dpdx = self.grad(p, deriv=(1, 0))
        dpdz = self.grad(p, deriv=(0, 1))
        u.vhat = u.vhat - cheby_to_galerkin(dpdx * fac, u)
        v.vhat = v.vhat - cheby_to_galerkin(dpdz * fac, v)

    def _phys(self, f):
        return f.dealias.backward(f.vhat) if self.dealias else f.backward(f.vhat)

    def update(self):
        """One adjoint-descent step (rbc2d_adj.py:321-356), replayed from a HIP
        graph after the first two calls when use_graph is set."""
        self._calls += 1
        if not self.use_graph or (self._graph is None and self._calls == 1):
            self._update_ops()
            return
        if self._graph is None:
            self._capture()
        else:
            # state the caller replaced since the capture goes into the captured buffers
            for f, buf in zip((self.U, self.V, self.T, self.pres), self._bufs):
                if f.vhat is not buf:
                    buf.copy_(ops.to_dev(f.vhat).reshape(buf.shape))
                    f.vhat = buf
        self._graph.replay()

    def _capture(self):
        """Capture one update.  U, V, T and pres carry over between updates: the
        captured update reads them from their current buffers and copies its
        results back into the same buffers, which the fields keep."""
        state = (self.U, self.V, self.T, self.pres)
        bufs = tuple(f.vhat for f in state)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        self._capturing = True
        try:
            with torch.cuda.graph(g):
                self._update_ops()
                for f, buf in zip(state, bufs):
                    if f.vhat is not buf:
                        buf.copy_(f.vhat)
        finally:
            self._capturing = False
            for f, buf in zip(state, bufs):
                f.vhat = buf
        self._graph, self._bufs = g, bufs

    def _update_ops(self):
        self.ux_old, self.uz_old, self.temp_old = 0, 0, 0
        for rk in range(self.nstage):
            self.ux, self.uz, self.temp = self._phys(self.U), self._phys(self.V), self._phys(self.T)
            self.update_NS()
            self.update_U(stage=rk)
            self.update_V(stage=rk)
            div = self.NS.divergence_velocity(self.U, self.V)
            self.update_P(div)
            self.update_pres(div, stage=rk)
            self.update_velocity(self.P, self.U, self.V)
            self.update_T(stage=rk)
            self.ux_old, self.uz_old, self.temp_old = self.ux, self.uz, self.temp

    def residual_norms(self):
        """(|R_U|, |R_V|, |R_T|) of the last residual, the norms the reference's
        callback prints (rbc2d_adj.py:302-319)."""
        return tuple(float(torch.linalg.vector_norm(f.vhat)) for f in (self.NS.U, self.NS.V,
                                                                        self.NS.T))

    def callback(self):
        nu = self.eval_Nu()
        div = self.NS.divergence_velocity(self.U, self.V)
        print("|div| = {:4.2e}".format(float(torch.linalg.vector_norm(div))))
        rU, rV, rT = self.residual_norms()
        print(" |U| = {:5.2e}".format(rU))
        print(" |V| = {:5.2e}".format(rV))
        print(" |T| = {:5.2e}".format(rT))
        return nu
