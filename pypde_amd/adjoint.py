"""Discrete adjoint of the fused RBC2D Euler step and gradients of trajectory
objectives with respect to the initial temperature.

The step of pypde_amd.engine.RBCBatch maps the state s = (X = [U, V, T], pres)
to s'.  It is linear except for the convective products, so its
vector-Jacobian product is the reverse sequence of the transposed operators:
transposed MFMA GEMMs, transposed four-diagonal and Poisson line solves
(pypde_fdma_solve_t, pypde_poisson_solve with transpose=1) and the adjoint of
the products (pypde_rbc_products_vjp), for which the two forward GEMMs that
produce the dealiased velocity and gradient fields are recomputed from the
step's input state.

Objectives (no reference counterpart; the reference's navier/rbc2d_adj.py:
321-356 is an adjoint-descent iteration toward steady states, a different
object, compared by direction in the tests):
  J_nu  = (1/n) sum_{k=1..n} Nu_plate(T_k)   time-averaged plate Nusselt number
  J_res = 0.5 || (Phi_dt(X) - X) / dt ||^2    steady-state residual (X = [T, U, V])
Long horizons use checkpoints every `ckpt` steps and recompute each segment
during the backward sweep.
"""
import ctypes

import numpy as np
import torch

from . import ops
from ._lib import check, lib
from .engine import AxisOps, RBCBatch, cheb_backward_matrix, cheb_forward_matrix, deriv_matrix, stencil


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _fields_gemm(ta, tb, m, n, k, batch, A, lda, iA, B, ldb, iB, C, ldc, iC, alpha=1.0,
                 alpha_vec=None, beta=0.0, D=None, iD=None, defer=False, alpha_div=3):
    """One batched GEMM over (instance, field) entries e = 3 b + f: operators
    indexed per field through a modulus-3 batch index, per-instance scales
    alpha_vec[e / alpha_div] (per instance: 3, per entry: 1); operands are
    device pointers.  defer: return the descriptor without launching."""
    d = ops._desc(ta, tb, m, n, k, batch, alpha, beta, A, lda, 0, B, ldb, 0, D,
                  ldc if D is not None else 0, 0, C, ldc, 0)
    d.iA, d.iB, d.iC = iA, iB, iC
    if D is not None:
        d.iD = iD
    if alpha_vec is not None:
        d.alpha_vec, d.alpha_div = alpha_vec.data_ptr(), alpha_div
    if defer:
        return d
    ops._launch_desc(d)


class StepAdjoint:
    """VJP of one fused step of an RBCBatch (all instances at once): Euler or
    the three-stage scheme, any implicit fraction beta.  A stage maps (X_s,
    pres_s, X_(s-1)) to (X_(s+1), pres_(s+1)) -- the previous stage's state
    enters through the velocities of the c-term (rbc2d.py:261-283) -- and the
    step's VJP runs the stage VJPs in reverse, adding each c-term cotangent to
    the stage before.

    The VJP runs on buffers of its own: the step input (X, pres) and the
    cotangents (Xb, pb, transformed in place) in `io`, every intermediate in
    one workspace arena, and for the three-stage scheme the engine's state
    and workspace, on which the stage states are recomputed.  Its launches
    are recorded once (after one op-by-op pass that tunes every kernel
    choice); vjp() replays them, and build_plan() records them into a step
    plan of the C-ABI (pypde_rbc_step_vjp)."""

    def __init__(self, eng: RBCBatch):
        self.e = eng
        e = eng
        B, N0, N1, M0, M1, L0, L1, K = e.B, e.N0, e.N1, e.M0, e.M1, e.L0, e.L1, e.K
        dev = e.X.device
        f64 = dict(dtype=ops.DTYPE, device=dev)
        NW = N1 + 2 * M1
        self.io = dict(X=torch.zeros((B, 3, M0, M1), **f64), pres=torch.zeros((B, N0, N1), **f64),
                       Xb=torch.zeros((B, 3, M0, M1), **f64), pb=torch.zeros((B, N0, N1), **f64))
        shapes = [("Wb", (B, M0, NW)), ("divb", (B, N0, N1)), ("Pb", (B, M0, M1)),
                  ("XPb", (B, M0, M1)), ("Q1b", (B, M0, N1)), ("Wdb", (B, 2 * M0, N1)),
                  ("Rb", (B, 3, M0, M1)), ("Rt", (3 * B, M1, M0)), ("Zb", (B, 3, K, M1)),
                  ("Prb", (B, 3, L0, L1)), ("Y1", (B, 3, 2 * L0, M1)),
                  ("Y2", (B, 3, 2 * L0, 2 * L1)), ("Y2b", (B, 3, 2 * L0, 2 * L1)),
                  ("y1b", (B, 3, 2 * L0, M1))]
        if e.nstage > 1:
            # the previous stage's fields and their cotangents (c-term), the
            # recomputed stage inputs, and the c-term cotangents (two, used in turn)
            shapes += [("Y2p", (B, 3, 2 * L0, 2 * L1)), ("Y2pb", (B, 3, 2 * L0, 2 * L1))]
            for st in range(1, e.nstage):
                shapes += [("Xs%d" % st, (B, 3, M0, M1)), ("ps%d" % st, (B, N0, N1))]
            shapes += [("cb0", (B, 3, M0, M1)), ("cb1", (B, 3, M0, M1))]
        total = sum(int(np.prod(sh)) for _, sh in shapes)
        self.ws = torch.zeros(total, **f64)
        self.v, at = {}, 0
        for name, sh in shapes:
            n = int(np.prod(sh))
            self.v[name] = self.ws[at:at + n].view(sh)
            at += n
        self.Y1, self.Y2, self.Y2b = self.v["Y1"], self.v["Y2"], self.v["Y2b"]
        self.t_mnub = torch.as_tensor(-e.nu * e.beta, **f64)
        self.graph = None
        self.launches = None
        self.plan = None
        # transposed solves on reversed lines (four-diagonal recurrences): the
        # Helmholtz tables per stage (their lambda is dt a_s beta nu), the
        # Poisson lines once
        self.hrt_st, self.hrt_y_st = [], []
        for st in range(e.nstage):
            hx, hy, prt, prt_lm = e.transposed_line_tables(st)
            self.hrt_st.append(hx)
            self.hrt_y_st.append(hy)
        self.hrt, self.hrt_y = self.hrt_st[0], self.hrt_y_st[0]
        self.prt = (prt, prt_lm)

    @staticmethod
    def _reversed_solve(x, n, nlines, batch, tabs, per_line, variant=None):
        """In place, along axis 0 of the contiguous x [batch, n, nlines]: the
        transposed line solves as the four-diagonal recurrence over each line
        read backwards (the first row at the last one, row stride -nlines).
        tabs: per-entry tables [batch][LM | R | MU | NU][n], or per-line tables
        ([n][nlines] slots, line-major slots).  The kernel form is tuned per
        shape on a scratch array; variant (1, 2, 3 or 4) forces a form."""
        from . import autotune
        from ._lib import Lines, Recurrence

        def geom(t):
            g = Lines()
            g.src = g.dst = t.data_ptr() + 8 * (n - 1) * nlines
            g.src_ld = g.dst_ld = -nlines
            g.src_bs = g.dst_bs = n * nlines
            g.len_in = g.len_out = n
            g.nlines, g.batch, g.axis, g.lines_per_block = nlines, batch, 0, 0
            return g

        def struct(v):
            r = Recurrence()
            if per_line:
                t = tabs[1] if v == 2 else tabs[0]
                r.LM, r.R, r.MU, r.NU = (t[i].data_ptr() for i in range(4))
                r.per_line, r.entry_stride = (2 if v == 2 else 1), 0
            else:
                p = tabs.data_ptr()
                r.LM, r.R, r.MU, r.NU = p, p + 8 * n, p + 16 * n, p + 24 * n
                r.per_line, r.entry_stride = 0, 4 * n
            r.F = r.pre = None
            r.variant = v
            return r

        def run(g, v):
            ops.call("transposed line solve", lib.pypde_recurrence_solve, ctypes.byref(g),
                     ctypes.byref(struct(v)))
        if variant is not None:
            run(geom(x), variant)
            return
        key = ("rec_reversed", n, nlines, batch, bool(per_line))
        if autotune.enabled() and not autotune._capturing() and key not in autotune.cache():
            scratch = torch.zeros_like(x)  # kept alive while the candidates run on it
            gs = geom(scratch)
            stream = ops.stream_ptr()
            autotune.choose(key, lambda v: check(lib.pypde_recurrence_solve(
                ctypes.byref(gs), ctypes.byref(struct(v)), stream), "transposed line solve"),
                (1, 2, 3, 4), default=2)
            torch.cuda.synchronize()
            del scratch
        run(geom(x), autotune.cache().get(key, 2))

    def forward_fields(self, X, Y2=None):
        """Y1 = GG_f X_f, Y2 = Y1 GGy^T (the dealiased values / gradients, into
        Y2 or the adjoint's own buffer), one launch each over every instance
        and field."""
        e = self.e
        Y2 = self.Y2 if Y2 is None else Y2
        B3, M0, M1, L0, L1 = 3 * e.B, e.M0, e.M1, e.L0, e.L1
        _fields_gemm(0, 0, 2 * L0, M1, M0, B3, e.GGx.data_ptr(), M0, ops._bi(2 * L0 * M0, m1=3),
                     X.data_ptr(), M1, ops._bi(M0 * M1), self.Y1.data_ptr(), M1,
                     ops._bi(2 * L0 * M1))
        _fields_gemm(0, 1, 2 * L0, 2 * L1, M1, B3, self.Y1.data_ptr(), M1, ops._bi(2 * L0 * M1),
                     e.GGy.data_ptr(), M1, ops._bi(0), Y2.data_ptr(), 2 * L1,
                     ops._bi(4 * L0 * L1))

    def fields_vjp(self, Y2b, out, accumulate):
        """out (+)= GG_f^T (Y2b GGy): the cotangent of X from that of its fields."""
        e = self.e
        B3, M0, M1, L0, L1 = 3 * e.B, e.M0, e.M1, e.L0, e.L1
        bi = ops._bi
        y1b = self.v["y1b"]
        _fields_gemm(0, 0, 2 * L0, M1, 2 * L1, B3, Y2b.data_ptr(), 2 * L1, bi(4 * L0 * L1),
                     e.GGy.data_ptr(), M1, bi(0), y1b.data_ptr(), M1, bi(2 * L0 * M1))
        _fields_gemm(1, 0, M0, M1, 2 * L0, B3, e.GGx.data_ptr(), M0, bi(2 * L0 * M0, m1=3),
                     y1b.data_ptr(), M1, bi(2 * L0 * M1), out.data_ptr(), M1, bi(M0 * M1),
                     beta=1.0 if accumulate else 0.0, D=out.data_ptr() if accumulate else None,
                     iD=bi(M0 * M1))

    # ---- the launch sequence ---------------------------------------------
    def _copy(self, src, dst):
        ops.call("copy", lib.pypde_axpby, src.numel(), 1.0, _vp(src), 0.0, _vp(dst))

    def _ops(self):
        """The step VJP op by op on the adjoint's buffers: io["Xb"], io["pb"]
        become the cotangents of io["X"], io["pres"].  For the three-stage scheme
        the stage inputs are first recomputed from (X, pres) by the engine's
        stage launches, on the engine's buffers."""
        e, io, v = self.e, self.io, self.v
        Xb, pb = io["Xb"], io["pb"]
        if e.nstage == 1:
            self._vjp_stage(0, io["X"], io["pres"], None, Xb, pb, None)
            return
        s = ops.stream_ptr()
        self._copy(io["X"], e.X)
        self._copy(io["pres"], e.pres)
        states = [(io["X"], io["pres"])]
        for st in range(e.nstage - 1):
            for f in e.launches:
                if e.stage_of(f) == st:
                    f(s)
                    ops.record(f)
            self._copy(e.X, v["Xs%d" % (st + 1)])
            self._copy(e.pres, v["ps%d" % (st + 1)])
            states.append((v["Xs%d" % (st + 1)], v["ps%d" % (st + 1)]))
        carry = None   # c-term cotangent of the stage input, from the stage after
        for st in range(e.nstage - 1, -1, -1):
            Xs, ps = states[st]
            Xprev = states[st - 1][0] if st > 0 else None
            cb = v["cb%d" % (st % 2)]
            has_c = self._vjp_stage(st, Xs, ps, Xprev, Xb, pb, cb)
            if carry is not None:
                ops.call("c-term cotangent", lib.pypde_axpby, Xb.numel(), 1.0, _vp(carry), 1.0,
                         _vp(Xb))
            carry = cb if has_c else None

    def _vjp_stage(self, st, X, pres, Xprev, Xb, pb, cb):
        """VJP of stage st, in place on the cotangents (Xb, pb) of its output,
        which become those of its input (X, pres); for a stage with a c-term,
        cb receives the cotangent of the previous stage's state Xprev (returns
        whether it did)."""
        e, v = self.e, self.v
        B, N0, N1, M0, M1, L0, L1, K = e.B, e.N0, e.N1, e.M0, e.M1, e.L0, e.L1, e.K
        g = ops.bgemm
        bi = ops._bi
        cterm = Xprev is not None and float(e.kc[st]) != 0.0
        # 21  velocity correction: W1b = -A_c0^T Ub, W2b = -A_c1^T Vb
        Wb = v["Wb"]
        ops.launch_grouped([
            g(e.Snx, pb, transA=True, alpha_vec=e.st_idta[st], out=Wb[:, :, :N1],
              defer=True)[1],  # 20
            g(e.Acorr[0], Xb[:, 0], transA=True, alpha=-1.0, out=Wb[:, :, N1:N1 + M1],
              defer=True)[1],
            g(e.Acorr[1], Xb[:, 1], transA=True, alpha=-1.0, out=Wb[:, :, N1 + M1:],
              defer=True)[1]], "vjp_correction")
        # 19  div^ = -nu beta pres^ (written whole: the axpby with beta 0 reads no divb)
        divb = v["divb"]
        ops.call("pres_div_vjp", lib.pypde_axpby, divb.numel(), 0.0, _vp(pb), 0.0, _vp(divb))
        ops.call("pres_div_vjp", lib.pypde_scaled_sub, N0 * N1, B, _vp(e.t_nu), e.beta, _vp(pb),
                 _vp(divb))
        # 18  P^ = W^ Wy
        Pb = g(Wb, e.Wy, out=v["Pb"])
        # 17  gauge
        ops.call("gauge_vjp", lib.pypde_set_entries, _vp(Pb), M0 * M1, B, 0.0)
        # 16  XP^ = P^ Qy ; 15 transposed Poisson line solves ; 14 Q1^ = RP^ Hy
        XPb = g(Pb, e.Qy, out=v["XPb"])
        self._reversed_solve(XPb, M0, M1, B, self.prt, True)
        Q1b = g(XPb, e.Hy, out=v["Q1b"])
        # 13  div^ += Bx^T Q1^
        g(e.Bx, Q1b, transA=True, beta=1.0, D=divb, out=divb)
        # 12  Wd^ = Cx^T div^   [B, 2M0, N1]
        Wdb = g(e.Cx, divb, transA=True, out=v["Wdb"])
        # 11  U*^ = Ub + Wd^_U C_y, V*^ = Vb + Wd^_V Cd_y, T*^ = Tb
        Rb = v["Rb"]
        ops.launch_grouped([
            g(Wdb[:, :M0], e.Cy[0], beta=1.0, D=Xb[:, 0], out=Rb[:, 0], defer=True)[1],
            g(Wdb[:, M0:], e.Cy[1], beta=1.0, D=Xb[:, 1], out=Rb[:, 1], defer=True)[1]],
            "vjp_div")
        # 9-10  transposed Helmholtz ADI: R^ = Hx^-T X*^ Hy^-1, both as
        # four-diagonal recurrences over reversed lines (tables of J H^T J);
        # along y on the transposed fields, so that the lines are columns (the
        # transposes take U*^, V*^ from Rb and T*^ = Tb from Xb)
        Rt = v["Rt"].view(B, 3, M1, M0)
        ops.call("vjp_transpose", lib.pypde_transpose, _vp(Rb), M1, 3 * M0 * M1, _vp(Rt), M0,
                 3 * M0 * M1, M0, M1, B)
        ops.call("vjp_transpose", lib.pypde_transpose, ctypes.c_void_p(Rb[:, 1].data_ptr()), M1,
                 3 * M0 * M1, ctypes.c_void_p(Rt[:, 1].data_ptr()), M0, 3 * M0 * M1, M0, M1, B)
        ops.call("vjp_transpose", lib.pypde_transpose, ctypes.c_void_p(Xb[:, 2].data_ptr()), M1,
                 3 * M0 * M1, ctypes.c_void_p(Rt[:, 2].data_ptr()), M0, 3 * M0 * M1, M0, M1, B)
        self._reversed_solve(v["Rt"], M1, M0, 3 * B, self.hrt_y_st[st], False)
        ops.call("vjp_transpose", lib.pypde_transpose, _vp(v["Rt"]), M0, M0 * M1, _vp(Rb), M1,
                 M0 * M1, M1, M0, 3 * B)
        R3 = Rb.view(3 * B, M0, M1)
        self._reversed_solve(R3, M0, M1, 3 * B, self.hrt_st[st], False)
        # 7  Z^_f = A_f^T R^_f ; 8 buoyancy Z^_T[2] += dt a BST_x^T R^_V
        Zb = v["Zb"]
        _fields_gemm(1, 0, K, M1, M0, 3 * B, e.Arhs.data_ptr(), K, bi(M0 * K, m1=3),
                     Rb.data_ptr(), M1, bi(M0 * M1), Zb.data_ptr(), M1, bi(K * M1))
        zt = Zb[:, 2, L0 + N0:L0 + N0 + M0]
        g(e.BSTx, Rb[:, 1], transA=True, alpha_vec=e.st_dta[st], beta=1.0, D=zt, out=zt)
        # 6  X^ (input) = Z^_f[2] BS_y  and  4  P^r_f = -dt b Z^_f[0] BF_y (one
        #    launch); the input cotangent overwrites Xb, no longer read
        Prb = v["Prb"]
        ops.launch_grouped([
            _fields_gemm(0, 0, M0, M1, M1, 3 * B, Zb.data_ptr() + 8 * (L0 + N0) * M1, M1,
                         bi(K * M1), e.BSy.data_ptr(), M1, bi(0), Xb.data_ptr(), M1,
                         bi(M0 * M1), defer=True),
            _fields_gemm(0, 0, L0, L1, M1, 3 * B, Zb.data_ptr(), M1, bi(K * M1),
                         e.BFy.data_ptr(), L1, bi(0), Prb.data_ptr(), L1, bi(L0 * L1),
                         alpha_vec=e.st_mdtb[st], defer=True)], "vjp_z")
        if e.diffusion:
            # 6b  explicit diffusion: Z[3] = f X BS_y^T, Z[4] = f X BD2S_y^T, so
            #     X^ += f (Z^[3] BS_y + Z^[4] BD2S_y), f per (instance, field)
            for blk in range(2):
                _fields_gemm(0, 0, M0, M1, M1, 3 * B,
                             Zb.data_ptr() + 8 * (L0 + N0 + (1 + blk) * M0) * M1, M1,
                             bi(K * M1), e.DFy[blk].data_ptr(), M1, bi(0), Xb.data_ptr(), M1,
                             bi(M0 * M1), alpha_vec=e.st_fdiff[st], alpha_div=1, beta=1.0,
                             D=Xb.data_ptr(), iD=bi(M0 * M1))
        # 5  pres^ (input) = pres^ - dt a (Z^_U[1] PBy_0 + Z^_V[1] PBy_1), in place
        g(Zb[:, 0, L0:L0 + N0], e.PBy[0], alpha_vec=e.st_mdta[st], beta=1.0, D=pb, out=pb)
        g(Zb[:, 1, L0:L0 + N0], e.PBy[1], alpha_vec=e.st_mdta[st], beta=1.0, D=pb, out=pb)
        # 3  products (with the c-term: the previous stage's velocities too)
        self.forward_fields(X)
        if cterm:
            self.forward_fields(Xprev, v["Y2p"])
            ratio = float(e.kc[st]) / float(e.kb[st])
            ops.call("products_vjp_c", lib.pypde_rbc_products_vjp_c, _vp(self.Y2), _vp(v["Y2p"]),
                     _vp(e.dTbcdz1), _vp(Prb), _vp(self.Y2b), _vp(v["Y2pb"]), ratio, L0, L1, B)
            self.fields_vjp(v["Y2pb"], cb, accumulate=False)
        else:
            ops.call("products_vjp", lib.pypde_rbc_products_vjp, _vp(self.Y2), _vp(e.dTbcdz1),
                     _vp(Prb), _vp(self.Y2b), L0, L1, B)
        # 2  Y1^ = Y2^ GGy ; 1 X^ += GG_f^T Y1^
        self.fields_vjp(self.Y2b, Xb, accumulate=True)
        return cterm

    def _build(self):
        """Record the launch sequence (after one op-by-op pass that tunes every
        kernel choice on the adjoint's buffers)."""
        if self.launches is None:
            saved = [t.clone() for t in (self.e.X, self.e.pres, self.e.P)]
            self._ops()
            with ops.record_launches() as L:
                self._ops()
            for t, s in zip((self.e.X, self.e.pres, self.e.P), saved):
                t.copy_(s)
            self.launches = L
        return self.launches

    def build_plan(self):
        """The step VJP as a step plan of the C-ABI: regions X, pres, Xb, pb,
        the adjoint's workspace and, for the three-stage scheme, the engine's
        X, pres, P and workspace, on which the stage states are recomputed
        (pypde_rbc_step_vjp carves the regions from 4 on out of one caller
        workspace)."""
        from .engine import Plan
        if self.plan is None:
            e, io = self.e, self.io
            regions = [(io[k].data_ptr(), 8 * io[k].numel()) for k in ("X", "pres", "Xb", "pb")]
            regions.append((self.ws.data_ptr(), 8 * self.ws.numel()))
            if e.nstage > 1:
                regions += [(t.data_ptr(), 8 * t.numel()) for t in (e.X, e.pres, e.P)]
                regions.append((e.ws.data_ptr(), e.workspace_bytes()))
            n = len(regions)
            bases = (ctypes.c_void_p * n)(*[b for b, _ in regions])
            sizes = (ctypes.c_longlong * n)(*[z for _, z in regions])
            plan = Plan(n, bases, sizes)
            for l in self._build():
                plan.add(l)
            self.plan = plan
        return self.plan

    def vjp(self, X, pres, Xb, pb):
        """Cotangents of the step input (X, pres) from those of its output (the
        recorded launches on the adjoint's buffers; new tensors).  The engine's
        own state is left as it was."""
        e, io = self.e, self.io
        self._build()
        saved = [t.clone() for t in (e.X, e.pres, e.P)] if e.nstage > 1 else None
        for k, t in zip(("X", "pres", "Xb", "pb"), (X, pres, Xb, pb)):
            io[k].copy_(t)
        s = ops.stream_ptr()
        for f in self.launches:
            f(s)
        if saved is not None:
            for t, v in zip((e.X, e.pres, e.P), saved):
                t.copy_(v)
        return io["Xb"].clone(), io["pb"].clone()

    def vjp_replay(self, X, pres, Xb, pb):
        """vjp() replayed from a HIP graph (captured on the first call, after one
        call that records the launches).  The same launches in the same order,
        so the results are vjp()'s bit for bit.  The returned tensors are the
        graph's output buffers: valid until the next call."""
        if self.graph is None:
            self.gin = [t.clone() for t in (X, pres, Xb, pb)]
            self.vjp(*self.gin)
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self.gout = self.vjp(*self.gin)
            self.graph = g
        for buf, t in zip(self.gin, (X, pres, Xb, pb)):
            buf.copy_(t)
        self.graph.replay()
        return self.gout


class NusseltFunctional:
    """Plate Nusselt number as an affine functional of the temperature
    coefficients: Nu(T) = c0 + a^T T b (rbc2d_base.py:344-363 is linear in T)."""

    def __init__(self, eng: RBCBatch):
        from .bases import Base
        e = eng
        N0, N1 = e.N0, e.N1
        Lz = 1.0  # y spans [-0.5, 0.5]
        x = Base(N0, "CH").x
        xm = np.zeros(N0 + 1)
        xm[0], xm[-1] = x[0], x[-1]
        xm[1:-1] = (x[1:] + x[:-1]) / 2.0
        dx = np.diff(xm)
        w = dx / dx.sum()
        Gx = cheb_backward_matrix(N0, N0) @ stencil(e.side, N0)     # T coefficients -> grid (x)
        Gy = cheb_backward_matrix(N1, N1) @ stencil("CD", N1)
        Dy = cheb_backward_matrix(N1, N1) @ deriv_matrix(N1) @ cheb_forward_matrix(N1, N1)
        ev = np.zeros(N1)
        ev[0] = ev[-1] = -0.5 * Lz
        scale = Lz / 2.0
        self.a = w @ Gx                                   # [M0]
        self.b = (ev @ (Dy @ Gy)) / scale                 # [M1]
        # constant part from the lifted boundary field Tbc, from its Chebyshev
        # coefficients by the FFT (the dense forward transform's rounding in the
        # high modes of this O(1) profile would reach the wall derivative)
        from .navier.rbc2d_base import cheb_coeffs_fft
        tbc = e.Tbc.v.detach().cpu().numpy()
        dy = cheb_coeffs_fft(tbc) @ deriv_matrix(N1).T / scale        # d/dy coefficients
        walls = dy @ np.stack([(-1.0) ** np.arange(N1), np.ones(N1)], axis=1)  # y = -1, +1
        vals = cheb_backward_matrix(N0, N0) @ walls                    # along x on the grid
        self.c0 = float(-0.5 * Lz * (w @ vals).sum())
        self.ta = torch.as_tensor(self.a, dtype=ops.DTYPE, device=e.X.device)
        self.tb = torch.as_tensor(self.b, dtype=ops.DTYPE, device=e.X.device)
        self.gmat = torch.outer(self.ta, self.tb)         # dNu/dT  [M0, M1]

    def value(self, T):
        """Nu for T [..., M0, M1] (one GEMM pair per call)."""
        tb = ops.bgemm(T if T.dim() == 3 else T.unsqueeze(0), self.tb.reshape(-1, 1))
        v = ops.bgemm(self.ta.reshape(1, -1), tb)
        return v.reshape(-1) + self.c0


class Trajectory:
    """Forward trajectory with checkpoints and its reverse (adjoint) sweep."""

    def __init__(self, eng: RBCBatch, ckpt=None, use_graph=True, budget=16 * 2**30):
        """ckpt: steps between stored states; None keeps every state when the
        trajectory fits in `budget` bytes of device memory (no recomputation),
        else spaces the checkpoints to fit."""
        self.e = eng
        self.adj = StepAdjoint(eng)
        self.nu = NusseltFunctional(eng)
        self.ckpt = ckpt
        self.budget = budget
        self.use_graph = use_graph

    def _spacing(self, nsteps):
        if self.ckpt is not None:
            return self.ckpt
        state = 8 * (self.e.X.numel() + self.e.pres.numel())
        return max(1, -(-nsteps * state // self.budget))

    def _step(self):
        """One forward step: the engine's graph replay, or its launches."""
        if self.use_graph:
            if 1 not in self.e.graphs:
                self.e.capture(1)
            self.e.step(1)
        else:
            self.e.step_eager()

    def _state(self):
        return self.e.X.clone(), self.e.pres.clone()

    def _set(self, X, pres):
        self.e.X.copy_(X)
        self.e.pres.copy_(pres)

    def run(self, X0, pres0, nsteps):
        """J_nu and the checkpoints; the engine state ends at step n."""
        self._set(X0, pres0)
        cks = {0: self._state()}
        acc = torch.zeros(self.e.B, dtype=ops.DTYPE, device=X0.device)
        ckpt = self._spacing(nsteps)
        for k in range(1, nsteps + 1):
            self._step()
            acc += self.nu.value(self.e.X[:, 2])
            if k % ckpt == 0 and k < nsteps:
                cks[k] = self._state()
        return acc / nsteps, cks

    def grad_nusselt(self, X0, pres0, nsteps):
        """dJ_nu / d(X0, pres0) by the reverse sweep over checkpointed segments."""
        J, cks = self.run(X0, pres0, nsteps)
        B = self.e.B
        Xb = torch.zeros_like(X0)
        pb = torch.zeros_like(pres0)
        g = self.nu.gmat / nsteps
        starts = sorted(cks)
        for si in range(len(starts) - 1, -1, -1):
            k0 = starts[si]
            k1 = starts[si + 1] if si + 1 < len(starts) else nsteps
            # recompute the segment's inputs X_k0 .. X_(k1-1) (the VJP of step k
            # needs X_(k-1); X_k1 itself is not needed)
            self._set(*cks[k0])
            states = []
            for k in range(k0, k1):
                states.append(self._state())
                if k + 1 < k1:
                    self._step()
            for k in range(k1, k0, -1):
                # Nu(T_k) enters J for k >= 1
                Xb[:, 2] += g
                X, pres = states[k - 1 - k0]
                if self.use_graph:
                    Xb, pb = self.adj.vjp_replay(X, pres, Xb, pb)
                else:
                    Xb, pb = self.adj.vjp(X, pres, Xb, pb)
        return J, Xb.clone(), pb.clone()


class TimeAveragedNusselt(torch.autograd.Function):
    """J = (1/n) sum_k Nu_plate(T_k) over n fused Euler steps from the initial
    temperature coefficients T0 (velocities and pressure fixed inputs);
    backward runs the checkpointed adjoint sweep."""

    @staticmethod
    def forward(ctx, T0, traj, X0, pres0, nsteps):
        X = X0.clone()
        X[:, 2] = T0
        J, _ = traj.run(X, pres0, nsteps)
        ctx.traj, ctx.nsteps = traj, nsteps
        ctx.save_for_backward(X, pres0)
        return J

    @staticmethod
    def backward(ctx, gJ):
        X, pres0 = ctx.saved_tensors
        _, Xb, _ = ctx.traj.grad_nusselt(X, pres0, ctx.nsteps)
        return Xb[:, 2] * gJ.reshape(-1, 1, 1), None, None, None, None


class RBCStep(torch.autograd.Function):
    """One fused Euler step (X, pres) -> (X', pres') with its hand-written VJP."""

    @staticmethod
    def forward(ctx, X, pres, adj):
        e = adj.e
        e.X.copy_(X)
        e.pres.copy_(pres)
        e.step_eager()
        ctx.adj = adj
        ctx.save_for_backward(X, pres)
        return e.X.clone(), e.pres.clone()

    @staticmethod
    def backward(ctx, Xb, pb):
        X, pres = ctx.saved_tensors
        Xb = torch.zeros_like(X) if Xb is None else Xb.contiguous()
        pb = torch.zeros_like(pres) if pb is None else pb.contiguous()
        Xin, pin = ctx.adj.vjp(X, pres, Xb, pb)
        return Xin, pin, None


def steady_residual(eng: RBCBatch, X, pres):
    """J_res = 0.5 ||(Phi(X) - X)/dt||^2 per instance over X = [U, V, T], pres
    held fixed (rbc2d_base.py:321-341 restated); the engine state is left at
    Phi(X)."""
    eng.X.copy_(X)
    eng.pres.copy_(pres)
    eng.step_eager()
    dt = torch.as_tensor(eng.dt, dtype=ops.DTYPE, device=X.device).reshape(-1, 1, 1, 1)
    r = (eng.X - X) / dt
    return 0.5 * (r * r).sum(dim=(1, 2, 3))


def steady_residual_grad(eng: RBCBatch, X, pres, adj=None):
    """J_res = 0.5 ||(Phi(X) - X)/dt||^2 over X = [U, V, T] with pres held
    fixed (rbc2d_base.py:321-341 restated), and dJ/dX."""
    adj = adj or StepAdjoint(eng)
    eng.X.copy_(X)
    eng.pres.copy_(pres)
    eng.step_eager()
    dt = torch.as_tensor(eng.dt, dtype=ops.DTYPE, device=X.device).reshape(-1, 1, 1, 1)
    r = (eng.X - X) / dt
    J = 0.5 * (r * r).sum(dim=(1, 2, 3))
    rb = r / dt
    Xin, _ = adj.vjp(X, pres, rb.contiguous(), torch.zeros_like(pres))
    return J, Xin - rb
