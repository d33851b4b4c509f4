"""Fused RBC2D time step on the GPU: a batch of same-shape instances with
per-instance Rayleigh / Prandtl numbers and time steps.

The Euler IMEX step of navier/rbc2d.py:396-434 is regrouped into 19 kernel
launches whatever the batch size: every separable 1-D operator chain of the
reference (stencil -> derivative -> zero-pad -> DCT, forward DCT -> truncate ->
pseudo-inverse B, ...) is folded into one dense operator per axis and applied
as a batched MFMA GEMM; the Helmholtz and Poisson solves stay banded line
solves with the reference's LU factors; the convective products and the
pressure update are pointwise kernels.

Data layout in HBM (float64, row-major):
  X    [B, 3, M0, M1]      Galerkin coefficients of U, V (ChebDirichlet^2) and
                           T (ChebNeumann or ChebDirichlet in x, ChebDirichlet in y)
  pres [B, N0, N1]         Chebyshev coefficients of the accumulated pressure
  P    [B, M0, M1]         pressure-correction potential (ChebNeumann^2)
with N = grid size, M = N-2, L = int(1.5 N) the dealiased grid.

Step (per instance; s = domain scale, dt/nu/kappa per instance):
   1  Y1_f = [G; Gd]_x(f) X_f                     x: values / x-derivatives on L0
   2  Y2_f = Y1_f [G; Gd]_y^T                     y: values / y-derivatives on L1
   3  P_f  = conv products (+ uz dTbc/dz for T)   pointwise on [L0, L1]
   4  Z_f[0]   = -dt P_f (B F)_y^T                 forward DCT + truncation + B, y
   5  Z_f[1]   = -dt pres (B | B D/s)_y^T          pressure gradient (U, V)
   6  Z_f[2]   = X_f (B S)_y^T                     old-value term
   7  X_f = [B F | B D/s or B | B S]_x Z_f + const_f   Helmholtz right-hand side
   8  X_V += dt (B S_T)_x Z_T[2]                   buoyancy
   9-10 Helmholtz ADI: four-diagonal LU sweeps along x then y (per-instance tables)
  11-12 div = (D S/s)_x U* S_y^T + S_x V* (D S/s)_y^T
  13-16 Poisson: B_x div H_y^T, per-line eigen-split solves, Q_y^T
  17    P[0,0] = 0 (in the epilogue of 16)
  18    W = P [S; T S; T D S/s]_y^T               (T = (S^T S)^-1 S^T)
  19-20 pres = pres - nu div beta + S_x W_0 / dt (one GEMM with a second addend)
  21    U = U* - (T D S/s)_x W_1,  V = V* - (T S)_x W_2
"""
import ctypes

import numpy as np
import torch

from . import ops
from ._lib import BIdx, GemmDesc, Lines, RbcRhs, check, lib
from .bases import Base, cheb_backward_matrix, cheb_forward_matrix, pseudoinverse_d2
from ._lib import Recurrence
from .solver import (fdma_lu, poisson_line_tables, rec_from_fdma, rec_from_poisson,
                     rec_poisson_reversed_transpose, rec_reversed_transpose)
from .templates import hholtz_matrices, poisson_matrices


# ---------------------------------------------------------------------------
# 1-D operator matrices (host, float64)
# ---------------------------------------------------------------------------
def deriv_matrix(N):
    """Dense Chebyshev-coefficient derivative, equal to the recurrence of
    differentiate_cheby.f90:28-53: (D c)_n = (2/c_n) sum_{p>n, p+n odd} p c_p."""
    D = np.zeros((N, N))
    for n in range(N):
        D[n, n + 1::2] = 2.0 * np.arange(n + 1, N, 2)
    D[0, :] *= 0.5
    return D


def stencil(key, N):
    return Base(N, key).stencil() if key != "CH" else np.eye(N)


def stencil_pinv(key, N):
    """(S^T S)^-1 S^T, the Galerkin projection of chebyshev.py:295-337."""
    S = stencil(key, N)
    return np.linalg.solve(S.T @ S, S.T)


def pinv_B(N):
    return pseudoinverse_d2(N)[2:, :]


class AxisOps:
    """All dense 1-D operators of one axis (grid size N, dealiased size L)."""

    def __init__(self, N, L, scale, side="CD"):
        self.N, self.M, self.L, self.s = N, N - 2, L, scale
        Bk = cheb_backward_matrix(L, N)
        F = cheb_forward_matrix(L, N)
        D1 = deriv_matrix(N)
        S = stencil("CD", N)
        St = stencil(side, N)
        Sn = stencil("CN", N)
        B = pinv_B(N)
        T = stencil_pinv("CD", N)
        self.G = Bk @ S                    # values on the L grid from CD coefficients
        self.Gd = Bk @ (D1 @ S) / scale    # derivative values
        self.GT = Bk @ St                  # the same for the temperature basis
        self.GdT = Bk @ (D1 @ St) / scale
        self.BF = B @ F                    # forward DCT, truncation to N, then B
        self.B = B
        self.BD = B @ D1 / scale           # pressure gradient, then B
        self.BS = B @ S
        self.BST = B @ St
        self.BD2S = B @ (D1 @ D1 @ S) / scale ** 2     # explicit diffusion (beta != 1)
        self.BD2ST = B @ (D1 @ D1 @ St) / scale ** 2
        self.C = S                         # divergence pieces
        self.Cd = D1 @ S / scale
        self.Sn = Sn                       # P (ChebNeumann) -> Chebyshev
        self.TSn = T @ Sn                  # velocity correction pieces
        self.TDSn = T @ D1 @ Sn / scale


# ---------------------------------------------------------------------------
# launch records
# ---------------------------------------------------------------------------
def _diags(A, offsets, tol=1e-9):
    """Diagonals of a banded operator: out[d, i] = A[i, i + offsets[d]] (0 where
    the column is outside A).  Entries off these diagonals must be rounding
    level (relative tol): the host products that form B D2 S leave up to about
    2e-10 of max|A| there at N = 512, which the banded form drops."""
    A = np.asarray(A)
    m, n = A.shape
    out = np.zeros((len(offsets), m))
    mask = np.zeros_like(A, dtype=bool)
    for d, off in enumerate(offsets):
        for i in range(m):
            if 0 <= i + off < n:
                out[d, i] = A[i, i + off]
                mask[i, i + off] = True
    rest = np.abs(np.where(mask, 0.0, A)).max() if A.size else 0.0
    if rest > tol * max(np.abs(A).max(), 1e-300):
        raise ValueError("operator is not banded on offsets %s (off-band %.3e)" % (offsets, rest))
    return out


def _bidx(s1=0, d1=1, m1=0, s2=0, d2=1, m2=0):
    b = BIdx()
    b.s1, b.d1, b.m1, b.s2, b.d2, b.m2 = s1, d1, m1, s2, d2, m2
    return b


def _mirror_signs(G, Lb, tol=1e-12):
    """Signs s_b with G[b Lb + Lb-1-j, k] = s_b (-1)^k G[b Lb + j, k] for every
    block b of Lb rows, or None when G is not mirror-symmetric this way."""
    nb = G.shape[0] // Lb
    par = (-1.0) ** np.arange(G.shape[1])
    signs = []
    for b in range(nb):
        blk = G[b * Lb:(b + 1) * Lb]
        scale = max(np.abs(blk).max(), 1e-300)
        found = None
        for sg in (1.0, -1.0):
            if np.abs(blk[::-1] - sg * par * blk).max() <= tol * scale:
                found = sg
                break
        if found is None:
            return None
        signs.append(found)
    return signs


def _sign_bits(signs):
    return sum(1 << b for b, sg in enumerate(signs) if sg < 0)


class _Gemm:
    def __init__(self, name, **kw):
        self.name = name
        d = GemmDesc()
        d.transA = kw.get("ta", 0)
        d.transB = kw.get("tb", 0)
        d.m, d.n, d.k, d.batch = kw["m"], kw["n"], kw["k"], kw["batch"]
        d.alpha = kw.get("alpha", 1.0)
        d.beta = kw.get("beta", 0.0)
        d.A, d.B, d.C = kw["A"], kw["B"], kw["C"]
        d.D = kw.get("D", None)
        d.lda, d.ldb, d.ldc = kw["lda"], kw["ldb"], kw["ldc"]
        d.ldd = kw.get("ldd", 0)
        d.iA = kw.get("iA", _bidx())
        d.iB = kw.get("iB", _bidx())
        d.iC = kw.get("iC", _bidx())
        d.iD = kw.get("iD", _bidx())
        d.alpha_vec = kw.get("alpha_vec", None)
        d.alpha_div = kw.get("alpha_div", 1)
        d.E = kw.get("E", None)
        d.lde = kw.get("lde", 0)
        d.iE = kw.get("iE", _bidx())
        d.gamma_vec = kw.get("gamma_vec", None)
        d.gamma = kw.get("gamma", 0.0)
        d.gamma_div = kw.get("gamma_div", 1)
        d.zero_first = kw.get("zero_first", 0)
        d.csn = kw.get("csn", 1)
        d.fold = kw.get("fold", 0)
        d.A1, d.B1 = kw.get("A1", None), kw.get("B1", None)
        d.lda1, d.ldb1, d.k1 = kw.get("lda1", 0), kw.get("ldb1", 0), kw.get("k1", 0)
        d.ska, d.skb = kw.get("ska", 1), kw.get("skb", 1)
        d.fold_len, d.fold_sign = kw.get("fold_len", 0), kw.get("fold_sign", 0)
        d.mirror_sign, d.mirror_div = kw.get("mirror_sign", 0.0), kw.get("mirror_div", 1)
        self.desc = d
        self.alt = None   # an equivalent unfolded _Gemm; the tuner keeps the faster
        # MFMA work actually issued (a fold halves it against the unfolded product)
        self.flops = 2.0 * d.m * d.n * d.k * d.batch * (2 if d.fold in (1, 2) else 1)

    def __call__(self, stream):
        check(lib.pypde_dgemm_batched(ctypes.byref(self.desc), stream), self.name)


class _GemmGroup:
    """Independent plain GEMMs of the step in one launch (pypde_dgemm_grouped):
    their tile grids are concatenated, so two latency-bound products share the
    device instead of running one after the other.  The tuner picks the
    group's kernel variant, or runs the members separately (split) when that
    is faster."""

    def __init__(self, members):
        self.members = list(members)
        self.name = "+".join(m.name for m in self.members)
        self.arr = (GemmDesc * len(self.members))()
        self.split = False
        self.flops = sum(m.flops for m in self.members)
        self.set_variant(self.members[0].desc.variant)

    @property
    def variant(self):
        return self.members[0].desc.variant

    def set_variant(self, v):
        for i, m in enumerate(self.members):
            m.desc.variant = v
            self.arr[i] = m.desc

    def __call__(self, stream):
        if self.split:
            for m in self.members:
                m(stream)
            return
        check(lib.pypde_dgemm_grouped(ctypes.byref(self.arr), len(self.members), stream),
              self.name)


class _SplitRhs:
    """Launch 7 with the right-hand-side product split along K: one launch of
    the partial products (batch over field x K slice), then
    pypde_splitk_reduce sums the slices in order onto the banded part.  For
    one or a few instances, where the unsplit product has too few output tiles
    to fill the device."""

    def __init__(self, name, group, reduce_args, S):
        self.name, self.group, self.reduce_args, self.S = name, group, reduce_args, S
        self.reduce = _Call("splitk_reduce", lib.pypde_splitk_reduce, *reduce_args)
        self.flops = group.flops

    def __call__(self, stream):
        self.group(stream)
        self.reduce(stream)


# independent launches of one stage that may share a launch
_GROUPS = (("ypass_v", "ypass_x"), ("div_x", "poisson_rhs"), ("poisson_qy", "corr_y", "corr_y0"),
           ("pres_p", "correct"))


def _group_launches(L, tag):
    by_name = {l.name: i for i, l in enumerate(L)}
    for names in _GROUPS:
        idx = [by_name.get(n + tag) for n in names]
        if any(i is None for i in idx):
            continue
        ms = [L[i] for i in idx]
        if not all(isinstance(m, _Gemm) and m.alt is None and m.desc.fold == 0 for m in ms):
            continue
        if len({(m.desc.transA, m.desc.transB) for m in ms}) != 1:
            continue
        if sorted(idx) != list(range(min(idx), min(idx) + len(idx))):
            continue  # only neighbours in the launch order
        L[min(idx)] = _GemmGroup(ms)
        for i in sorted(idx)[1:][::-1]:
            del L[i]
        by_name = {l.name: i for i, l in enumerate(L)}


class _Call:
    def __init__(self, name, fn, *args):
        self.name, self.fn, self.args = name, fn, args

    def __call__(self, stream):
        check(self.fn(*self.args, stream), self.name)


class Plan:
    """A step plan of the C-ABI (pypde_plan_create ... pypde_plan_destroy):
    launches recorded with pointers relative to the given regions."""

    # the recording entry point of each single-kernel launch of the step
    _ADD = {"pypde_rbc_products": "pypde_plan_add_products",
            "pypde_rbc_products_yfold": "pypde_plan_add_products_yfold",
            "pypde_splitk_reduce": "pypde_plan_add_splitk_reduce",
            "pypde_rbc_rhs_banded": "pypde_plan_add_rhs_banded",
            "pypde_rbc_products_rhs": "pypde_plan_add_products_rhs",
            "pypde_fdma_solve": "pypde_plan_add_fdma",
            "pypde_poisson_solve": "pypde_plan_add_poisson",
            "pypde_recurrence_solve": "pypde_plan_add_recurrence",
            "pypde_axpby": "pypde_plan_add_axpby",
            "pypde_transpose": "pypde_plan_add_transpose",
            "pypde_scaled_sub": "pypde_plan_add_scaled_sub",
            "pypde_set_entries": "pypde_plan_add_set_entries",
            "pypde_rbc_products_vjp": "pypde_plan_add_products_vjp",
            "pypde_rbc_products_vjp_c": "pypde_plan_add_products_vjp_c"}

    def __init__(self, nregions, bases, sizes):
        h = ctypes.c_void_p()
        check(lib.pypde_plan_create(nregions, bases, sizes, ctypes.byref(h)), "pypde_plan_create")
        self.handle = h

    def add(self, launch):
        """Record one launch (a _Gemm, _GemmGroup, _RecCall, _Call, _SplitRhs or
        an ops.Recorded)."""
        h = self.handle
        if isinstance(launch, _SplitRhs):
            self.add(launch.group)
            self.add(launch.reduce)
        elif isinstance(launch, ops.Recorded):
            if launch.kind == "gemm":
                check(lib.pypde_plan_add_gemm(h, ctypes.byref(launch.arr)), launch.name)
            elif launch.kind == "group":
                check(lib.pypde_plan_add_grouped(h, ctypes.byref(launch.arr), launch.n),
                      launch.name)
            elif launch.fn.__name__ in Plan._ADD:
                check(getattr(lib, Plan._ADD[launch.fn.__name__])(h, *launch.args), launch.name)
            else:
                raise TypeError("launch %r cannot be recorded in a step plan" % launch.name)
        elif isinstance(launch, _Gemm):
            check(lib.pypde_plan_add_gemm(h, ctypes.byref(launch.desc)), launch.name)
        elif isinstance(launch, _GemmGroup):
            if launch.split:
                for m in launch.members:
                    check(lib.pypde_plan_add_gemm(h, ctypes.byref(m.desc)), m.name)
            else:
                check(lib.pypde_plan_add_grouped(h, ctypes.byref(launch.arr), len(launch.members)),
                      launch.name)
        elif isinstance(launch, _RecCall):
            check(lib.pypde_plan_add_recurrence(h, *launch.args), launch.name)
        elif isinstance(launch, _Call) and launch.fn.__name__ in Plan._ADD:
            add = getattr(lib, Plan._ADD[launch.fn.__name__])
            check(add(h, *launch.args), launch.name)
        else:
            raise TypeError("launch %r cannot be recorded in a step plan" % (launch,))

    def size(self):
        return lib.pypde_plan_size(self.handle)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and lib is not None:
            try:
                lib.pypde_plan_destroy(h)
            except Exception:
                pass
            self.handle = None


# recurrence kernel forms: 1 block (64 lines per block), 2 lane-chunk,
# 3 block with 32 lines per block, 4 the same with 32 positions per chunk
REC_VARIANTS = (1, 2, 3, 4)


class _RecCall(_Call):
    """A recurrence launch with a struct per kernel form (REC_VARIANTS); the
    engine's tuner picks one."""

    def __init__(self, name, g, structs, variant):
        self.g, self.structs = g, structs
        _Call.__init__(self, name, lib.pypde_recurrence_solve, ctypes.byref(g),
                       ctypes.byref(structs[variant]))
        self.variant = variant

    def set_variant(self, v):
        self.variant = v
        self.args = (ctypes.byref(self.g), ctypes.byref(self.structs[v]))

    def key(self):
        g, r = self.g, self.structs[1]
        return ("rec", g.axis, g.len_in, g.len_out, g.nlines, g.batch, bool(r.per_line),
                r.F is not None, r.pre is not None)


def _rec_default(lines):
    """Lane-chunk form for few lines, block form otherwise."""
    return 2 if lines <= 2048 else 1


def _lines(src, dst, ld, bs, n, nlines, batch, axis):
    g = Lines()
    g.src, g.dst = src, dst
    g.src_ld = g.dst_ld = ld
    g.src_bs = g.dst_bs = bs
    g.len_in = g.len_out = n
    g.nlines, g.batch, g.axis = nlines, batch, axis
    g.lines_per_block = 0
    return g


def _rec(base, M, entry_stride, variant=0):
    """Recurrence tables stacked [entry][LM | R | MU | NU][M] at `base`."""
    r = Recurrence()
    r.LM, r.R, r.MU, r.NU = base, base + 8 * M, base + 16 * M, base + 24 * M
    r.F = r.pre = None
    r.entry_stride, r.per_line, r.variant = entry_stride, 0, variant
    return r


# ---------------------------------------------------------------------------
class RBCBatch:
    """IMEX RBC2D step (Euler, or three-stage RK3) for `B` instances of one shape.

    params: list of dicts with keys ra, pr, dt (one per instance).
    case: "rbc" (adiabatic sidewalls, T in ChebNeumann x ChebDirichlet),
          "linear" or "zero" (T in ChebDirichlet^2).
    integrator: "eu" or "rk3" (the stage coefficients a, b, c of
          rbc2d_base.py:86-102); beta: implicit fraction of the diffusion, the
          rest treated explicitly (rbc2d.py:269-283).
    group_launches: independent GEMMs of a stage share one launch
          (pypde_dgemm_grouped) where the tuner finds it faster, and for few
          instances the right-hand-side product may be split along K.
    """

    def __init__(self, shape, params, case="rbc", dealias=True, aspect=1.0, beta=1.0,
                 integrator="eu", group_launches=True, exact_solves=False):
        self.group_launches = bool(group_launches)
        # the Helmholtz and Poisson line solves as the LDS-panel kernels that
        # keep the reference's LU statement order (pypde_fdma_solve,
        # pypde_poisson_solve), instead of the faster register recurrences
        self.exact_solves = bool(exact_solves)
        if case not in ("rbc", "linear", "zero"):
            raise ValueError("Specified case is not available: ", ["rbc", "linear", "zero"])
        if integrator not in ("eu", "rk3"):
            raise ValueError("Specified integrator is not available: ", ["eu", "rk3"])
        self.integrator = integrator
        if integrator == "rk3":
            self.ka = np.array([8.0 / 15.0, 2.0 / 15.0, 1.0 / 3.0])
            self.kb = np.array([8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0])
            self.kc = np.array([0, -17.0 / 60.0, -5.0 / 12.0])
        else:
            self.ka, self.kb, self.kc = np.array([1.0]), np.array([1.0]), np.array([0])
        self.nstage = len(self.ka)
        self.shape = tuple(shape)
        self.case, self.dealias, self.aspect, self.beta = case, dealias, aspect, float(beta)
        self.params = [dict(p) for p in params]
        B = self.B = len(params)
        N0, N1 = self.shape
        M0, M1 = N0 - 2, N1 - 2
        L0 = int(N0 * 1.5) if dealias else N0
        L1 = int(N1 * 1.5) if dealias else N1
        self.N0, self.N1, self.M0, self.M1, self.L0, self.L1 = N0, N1, M0, M1, L0, L1
        self.scale = (aspect * 0.5, 0.5)
        side = "CN" if case == "rbc" else "CD"
        self.side = side
        ax = AxisOps(N0, L0, self.scale[0], side)
        ay = AxisOps(N1, L1, self.scale[1], "CD")
        self.ax, self.ay = ax, ay
        dev = ops.device()
        f64 = dict(dtype=ops.DTYPE, device=dev)
        T = lambda a: torch.as_tensor(np.ascontiguousarray(a), **f64)

        # per-instance scalars
        ra = np.array([p["ra"] for p in params], float)
        pr = np.array([p.get("pr", 1.0) for p in params], float)
        dt = np.array([p["dt"] for p in params], float)
        # nu, kappa from Ra and Pr over a unit height (rbc2d_base.py:46-56)
        # unless the instance gives them
        nu = np.array([p["nu"] if "nu" in p else np.sqrt(b / (a / 1.0 ** 3.0))
                       for p, a, b in zip(params, ra, pr)], float)
        kappa = np.array([p["kappa"] if "kappa" in p else np.sqrt(1 / b / (a / 1.0 ** 3.0))
                          for p, a, b in zip(params, ra, pr)], float)
        self.dt, self.nu, self.kappa = dt, nu, kappa
        self.t_mdt = T(-dt)
        self.t_dt = T(dt)
        self.t_idt = T(1.0 / dt)
        self.t_nu = T(nu)
        # per stage: -dt b (convection), -dt a (pressure gradient), dt a (buoyancy),
        # 1 / (dt a) (pressure update), explicit diffusion dt a (1 - beta) coef
        self.st_mdtb = [T(-dt * self.kb[s]) for s in range(self.nstage)]
        self.st_mdta = [T(-dt * self.ka[s]) for s in range(self.nstage)]
        self.st_dta = [T(dt * self.ka[s]) for s in range(self.nstage)]
        self.st_idta = [T(1.0 / (dt * self.ka[s])) for s in range(self.nstage)]
        coef = np.stack([nu, nu, kappa], axis=1)                               # [B, 3]
        self.st_fdiff = [T((dt * self.ka[s] * (1 - self.beta))[:, None] * coef)
                         for s in range(self.nstage)]
        self.diffusion = self.beta != 1.0

        # ---- operator stacks (device)
        self.GGx = T(np.stack([np.vstack([ax.G, ax.Gd]), np.vstack([ax.G, ax.Gd]),
                               np.vstack([ax.GT, ax.GdT])]))                   # [3, 2L0, M0]
        self.GGy = T(np.vstack([ay.G, ay.Gd]))                                # [2L1, M1]
        self.BFy = T(ay.BF)                                                   # [M1, L1]
        # parity-folded halves of the transform operators (even lengths only)
        self.fold_x = self.fold_y = self.fold_bf = None
        if M0 % 2 == 0 and L0 % 2 == 0:
            GG = [np.vstack([ax.G, ax.Gd]), np.vstack([ax.GT, ax.GdT])]
            sg = [_mirror_signs(g, L0) for g in GG]
            if sg[0] is not None and sg[0] == sg[1]:
                tops = [np.vstack([g[:L0 // 2], g[L0:L0 + L0 // 2]]) for g in (GG[0], GG[0], GG[1])]
                self.fold_x = dict(E=T(np.stack([t[:, 0::2] for t in tops])),
                                   O=T(np.stack([t[:, 1::2] for t in tops])), sign=_sign_bits(sg[0]))
        if M1 % 2 == 0 and L1 % 2 == 0:
            GGy_h = np.vstack([ay.G, ay.Gd])
            sg = _mirror_signs(GGy_h, L1)
            if sg is not None:
                top = np.vstack([GGy_h[:L1 // 2], GGy_h[L1:L1 + L1 // 2]])
                self.fold_y = dict(E=T(top[:, 0::2]), O=T(top[:, 1::2]), sign=_sign_bits(sg))
            sg = _mirror_signs(ay.BF.T, L1)   # BF[i, L1-1-j] = s (-1)^i BF[i, j]
            if sg is not None:
                h = L1 // 2
                self.fold_bf = dict(B=T(np.stack([ay.BF[0::2, :h], ay.BF[1::2, :h]])), sign=sg[0])
        # y-folded layout of the dealiased fields: the y-pass writes even / odd
        # coefficient halves over the first half of the y grid and the products
        # kernel reconstructs both mirrored columns (half-size plain GEMMs)
        self.yfold = None
        if self.fold_y is not None and self.fold_bf is not None:
            sv = _mirror_signs(ay.G, L1)
            sd = _mirror_signs(ay.Gd, L1)
            if sv is not None and sd is not None:
                H = L1 // 2
                ops_y = np.stack([ay.G[:H, 0::2], ay.G[:H, 1::2], ay.Gd[:H, 0::2], ay.Gd[:H, 1::2]])
                self.yfold = dict(B=T(ops_y), sv=sv[0], sd=sd[0])
        self.PBy = T(np.stack([ay.B, ay.BD]))                                 # [2, M1, N1]
        self.BSy = T(ay.BS)                                                   # [M1, M1]
        K = L0 + N0 + M0 + (2 * M0 if self.beta != 1.0 else 0)
        self.K = K
        zero_blk = np.zeros((M0, N0))
        blocks = [[ax.BF, ax.BD, ax.BS], [ax.BF, ax.B, ax.BS], [ax.BF, zero_blk, ax.BST]]
        if self.beta != 1.0:
            # explicit diffusion f (d2x + d2y) of the stage-start field, applied
            # through B: (B D2 S)_x X (B S)_y^T + (B S)_x X (B D2 S)_y^T
            blocks[0] += [ax.BD2S, ax.BS]
            blocks[1] += [ax.BD2S, ax.BS]
            blocks[2] += [ax.BD2ST, ax.BST]
            self.DFy = T(np.stack([ay.BS, ay.BD2S]))                          # [2, M1, M1]
        # the dense right-hand-side operator of the discrete adjoint (its
        # transpose maps the cotangent of X to those of the stacked parts)
        self.Arhs = T(np.stack([np.hstack(b) for b in blocks]))              # [3, M0, K]
        self.BSTx = T(ax.BST)                                                 # [M0, M0]
        # the forward step applies the banded blocks as diagonals
        # (pypde_rbc_rhs_banded) and only the forward transform BF densely
        self.BFx = T(ax.BF)                                                   # [M0, L0]
        self.tx = T(np.vstack([_diags(ax.BD, (1, 3, 5)), _diags(ax.B, (0, 2, 4)),
                               _diags(ax.BS, (-2, 0, 2, 4)), _diags(ax.BST, (-2, 0, 2, 4)),
                               _diags(ax.BD2S, (0, 2)), _diags(ax.BD2ST, (0, 2))]))   # [18, M0]
        self.ty = T(np.vstack([_diags(ay.B, (0, 2, 4)), _diags(ay.BD, (1, 3, 5)),
                               _diags(ay.BS, (-2, 0, 2, 4)), _diags(ay.BD2S, (0, 2))]))  # [12, M1]
        self.Cy = T(np.stack([ay.C, ay.Cd]))                                  # [2, N1, M1]
        self.Cx = T(np.hstack([ax.Cd, ax.C]))                                 # [N0, 2M0]
        pm = poisson_matrices([Base(N0, "CN"), Base(N1, "CN")], singular=True, scale=self.scale)
        self.Bx = T(ax.B)                                                     # [M0, N0]
        self.Hy = T(pm["Hy"])                                                 # [M1, N1]
        self.Qy = T(pm["Qy"])                                                 # [M1, M1]
        tab, sing = poisson_line_tables(pm["Ax"], pm["Cx"], pm["wy"], True)
        self.ptab = T(tab)
        self.psing = torch.as_tensor(sing, device=dev)
        self.prec = T(np.stack(rec_from_poisson(tab, sing)))  # [LM | R | MU | NU][M0][M1]
        self.precT = self.prec.transpose(1, 2).contiguous()   # line-major [4][M1][M0]
        Wy = np.vstack([ay.Sn, ay.TSn, ay.TDSn])
        self.Wy = T(Wy)                                                       # [N1+2M1, M1]
        # composites that take the Poisson chain off the critical path:
        #   RP = B_x div H_y^T = (B [Cd | C])_x [X_f (H_y C_f)^T]_f
        #   W1 = P W_y^T = RP (W_y Q_y)^T, row 0 with row 0 of Q_y zeroed (the
        #        gauge P[0,0] = 0 of the solve)
        self.HCy = T(np.stack([np.vstack([ay.C, pm["Hy"] @ ay.C]),
                               np.vstack([ay.Cd, pm["Hy"] @ ay.Cd])]))        # [2, N1+M1, M1]
        self.BCx = T(ax.B @ np.hstack([ax.Cd, ax.C]))                         # [M0, 2M0]
        Qy0 = np.array(pm["Qy"], copy=True)
        Qy0[0, :] = 0.0
        self.WQy = T(Wy @ pm["Qy"])                                           # [N1+2M1, M1]
        self.WQy0 = T(Wy @ Qy0)
        self.Snx = T(ax.Sn)                                                   # [N0, M0]
        self.Acorr = T(np.stack([ax.TDSn, ax.TSn]))                           # [2, M0, M0]

        # Helmholtz LU tables per stage, instance and field: [B, 3, 4, M]
        self.st_hx, self.st_hy, self.st_rhx, self.st_rhy = [], [], [], []
        for st in range(self.nstage):
            hx = np.zeros((B, 3, 4, M0))
            hy = np.zeros((B, 3, 4, M1))
            for b in range(B):
                for f in range(3):
                    hx[b, f] = self._lu(self._helmholtz_matrix(st, b, f, 0))
                    hy[b, f] = self._lu(self._helmholtz_matrix(st, b, f, 1))
            self.st_hx.append(T(hx))
            self.st_hy.append(T(hy))
            # normalised (reciprocal-pivot) form for the register-resident kernel
            self.st_rhx.append(T(np.stack(rec_from_fdma(hx[:, :, 0, :M0 - 2], hx[:, :, 1],
                                                        hx[:, :, 2, :M0 - 2], hx[:, :, 3, :M0 - 4]),
                                          axis=2)))
            self.st_rhy.append(T(np.stack(rec_from_fdma(hy[:, :, 0, :M1 - 2], hy[:, :, 1],
                                                        hy[:, :, 2, :M1 - 2], hy[:, :, 3, :M1 - 4]),
                                          axis=2)))
        self.hx, self.hy, self.rhx, self.rhy = (self.st_hx[0], self.st_hy[0], self.st_rhx[0],
                                                self.st_rhy[0])

        # boundary lifting constants (shape-only) and per-instance RHS constants
        self._setup_bc()
        Bx_, By_ = ax.B, ay.B
        cv = Bx_ @ self.Tbc_cheby_h @ By_.T
        ct = Bx_ @ self.dTbcdz2_h @ By_.T
        self.st_const = []
        for st in range(self.nstage):
            cst = np.zeros((B, 3, M0, M1))
            for b in range(B):
                cst[b, 1] = dt[b] * self.ka[st] * cv
                cst[b, 2] = dt[b] * self.ka[st] * kappa[b] * ct
            self.st_const.append(T(cst))
        self.const = self.st_const[0]

        # ---- state and workspaces
        self.X = torch.zeros((B, 3, M0, M1), **f64)
        self.pres = torch.zeros((B, N0, N1), **f64)
        self.P = torch.zeros((B, M0, M1), **f64)
        if self.yfold is None and (self.nstage > 1 or self.beta != 1.0):
            raise NotImplementedError("RK3 and beta != 1 need even dealiased lengths in y "
                                      "(the y-folded layout)")
        # candidate K splits of the right-hand-side product (few instances only)
        tiles = 3 * B * ((M0 + 31) // 32) * ((M1 + 31) // 32)
        self._splits = [S for S in (2, 3, 4, 6, 8) if L0 % S == 0 and L0 // S >= 32] \
            if (tiles <= 768 and self.group_launches) else []
        # every workspace of the step is a view into one buffer (the step plan's
        # workspace region); the K-split partials come last, so that a plan
        # that keeps fewer slices needs a shorter workspace
        wsd = [("Y1", (B, 3, 2 * L0, M1))]
        if self.yfold is not None:
            wsd.append(("Y2", (B, 3, 6, L0, L1 // 2)))
            if self.nstage > 1:  # the previous RK3 stage's velocities (c-term)
                wsd.append(("Y2old", (B, 3, 6, L0, L1 // 2)))
            wsd.append(("Pr", (B, 3, 2, L0, L1 // 2)))
        else:
            wsd += [("Y2", (B, 3, 2 * L0, 2 * L1)), ("Pr", (B, 3, L0, L1))]
        wsd += [("Z", (B, 3, L0, M1)), ("Rb", (B, 3, M0, M1)), ("div", (B, N0, N1)),
                ("W2", (B, 2, M0, N1 + M1)), ("RP", (B, M0, M1)), ("W1", (B, M0, N1 + 2 * M1))]
        if self._splits:
            wsd += [("rhs_work", (3 * B * max(self._splits), M0, M1))]
        offs, total = {}, 0
        for name, shape in wsd:
            offs[name] = total
            total += -(-int(np.prod(shape)) // 32) * 32   # 256-byte aligned views
        self.ws = torch.zeros(total, **f64)
        self.ws_offsets = offs
        for name, shape in wsd:
            n = int(np.prod(shape))
            setattr(self, name, self.ws[offs[name]:offs[name] + n].view(shape))
        if self.nstage == 1:
            self.Y2old = None
        self._build_launches()
        self.tune()
        self.graph = None
        self.graph_steps = 0
        self.graphs = {}  # captured graphs by their number of steps
        self.plan = None

    def _helmholtz_matrix(self, st, b, f, axis):
        """Dense Helmholtz operator B S - lam I S of stage st, instance b, field
        f (U, V, T) along axis (hholtz.py:70-93): lam = dt a beta nu (U, V) or
        dt a beta kappa (T); T's x basis follows the case."""
        key = self.side if (axis == 0 and f == 2) else "CD"
        n = self.N0 if axis == 0 else self.N1
        coef = self.kappa[b] if f == 2 else self.nu[b]
        lam = self.dt[b] * self.ka[st] * self.beta * coef
        bases = self.__dict__.setdefault("_hbases", {})
        if (n, key) not in bases:
            bases[(n, key)] = Base(n, key)
        return hholtz_matrices(bases[(n, key)], lam, self.scale[axis])[2]

    def transposed_line_tables(self, st=0):
        """Normalised tables of the transposed line solves on reversed lines
        (stage st), for the discrete adjoint: the Helmholtz solves along x and
        along y per (instance, field) entry, [B, 3, LM | R | MU | NU, M], and the
        Poisson line solves per line, [LM | R | MU | NU][M0][M1] and line-major
        [4][M1][M0]."""
        f64 = dict(dtype=ops.DTYPE, device=self.X.device)
        hx = np.zeros((self.B, 3, 4, self.M0))
        hy = np.zeros((self.B, 3, 4, self.M1))
        for b in range(self.B):
            for f in range(3):
                hx[b, f] = np.stack(rec_reversed_transpose(self._helmholtz_matrix(st, b, f, 0)))
                hy[b, f] = np.stack(rec_reversed_transpose(self._helmholtz_matrix(st, b, f, 1)))
        pm = poisson_matrices([Base(self.N0, "CN"), Base(self.N1, "CN")], singular=True,
                              scale=self.scale)
        pr = torch.as_tensor(np.stack(rec_poisson_reversed_transpose(pm["Ax"], pm["Cx"],
                                                                     pm["wy"], True)), **f64)
        return (torch.as_tensor(hx, **f64), torch.as_tensor(hy, **f64), pr,
                pr.transpose(1, 2).contiguous())

    @staticmethod
    def _lu(A):
        A = np.asarray(A)
        l, d = np.diag(A, -2).copy(), np.diag(A).copy()
        u1, u2 = np.diag(A, 2).copy(), np.diag(A, 4).copy()
        fdma_lu(l, d, u1, u2)
        M = d.shape[0]
        out = np.zeros((4, M))
        out[0, :M - 2], out[1], out[2, :M - 2], out[3, :M - 4] = l, d, u1, u2
        return out

    def _setup_bc(self):
        """Lifted temperature boundary field and its derivatives
        (rbc2d.py:135-178), computed with the drop-in field operations."""
        from .field import Field, FieldBC
        from .field_operations import galerkin_to_cheby, grad
        N0, N1 = self.shape
        xs = [Base(N0, self.side, dealias=1.5), Base(N1, "CD", dealias=1.5)]
        deriv_field = Field([Base(N0, "CH", dealias=1.5), Base(N1, "CH", dealias=1.5)])
        if self.case == "zero":
            y = deriv_field.y * self.scale[1]
            bc = np.zeros((2, N1))
            bc[0, :] = transfer_function(0.5, 0, -0.5, y, k=0.02)
            bc[1, :] = bc[0, :]
            Tbc = FieldBC(xs, axis=0)
        else:
            bc = np.zeros((N0, 2))
            bc[:, 0], bc[:, 1] = 0.5, -0.5
            Tbc = FieldBC(xs, axis=1)
        Tbc.add_bc(bc)
        self.Tbc = Tbc
        self.dTbcdz2 = grad(Tbc, (0, 2), scale=self.scale)
        vhat = grad(Tbc, (0, 1), scale=self.scale)
        if self.dealias:
            self.dTbcdz1 = deriv_field.dealias.backward(vhat).contiguous()
        else:
            self.dTbcdz1 = deriv_field.backward(vhat).contiguous()
        self.Tbc_cheby = galerkin_to_cheby(Tbc.vhat, Tbc)
        self.Tbc_cheby_h = self.Tbc_cheby.cpu().numpy()
        self.dTbcdz2_h = self.dTbcdz2.cpu().numpy()
        self.deriv_field = deriv_field

    # -----------------------------------------------------------------------
    def _build_launches(self):
        self._keep = []
        self._rhs_pairs = []
        L = []
        for st in range(self.nstage):
            S = []
            self._stage_launches(S, st)
            if self.group_launches:
                _group_launches(S, "" if self.nstage == 1 else "[%d]" % st)
            L += S
        self.launches = L

    def _stage_launches(self, L, st):
        """The launches of stage `st` (one stage for Euler, three for RK3)."""
        B, N0, N1, M0, M1, L0, L1, K = (self.B, self.N0, self.N1, self.M0, self.M1, self.L0,
                                        self.L1, self.K)
        p = lambda t: t.data_ptr()
        X, Y1, Pr, Z = self.X, self.Y1, self.Pr, self.Z
        sX = M0 * M1
        sZ = self.sZ = L0 * M1
        tag = "" if self.nstage == 1 else "[%d]" % st
        # 1  x-pass: values and x-derivatives on the L0 grid
        g = _Gemm("xpass" + tag, m=2 * L0, n=M1, k=M0, batch=3 * B, A=p(self.GGx), lda=M0,
                  iA=_bidx(s1=2 * L0 * M0, m1=3), B=p(X), ldb=M1, iB=_bidx(s1=sX),
                  C=p(Y1), ldc=M1, iC=_bidx(s1=2 * L0 * M1))
        if self.fold_x is not None:
            fx, h = self.fold_x, M0 // 2
            g = self._with_alt(g, _Gemm(
                "xpass" + tag, m=L0, n=M1, k=h, batch=3 * B, A=p(fx["E"]), lda=h,
                iA=_bidx(s1=L0 * h, m1=3), B=p(X), ldb=2 * M1, iB=_bidx(s1=sX),
                C=p(Y1), ldc=M1, iC=_bidx(s1=2 * L0 * M1), fold=1, A1=p(fx["O"]),
                B1=p(X) + 8 * M1, lda1=h, ldb1=2 * M1, k1=h, fold_len=L0, fold_sign=fx["sign"]))
        L.append(g)
        if self.yfold is not None:
            self._yfold_launches(L, st, tag)
        else:
            self._plain_y_launches(L, tag)
        # 5-6  banded part R_f of the right-hand side from the stage-start state:
        #      constant, pressure gradient, old value, buoyancy, explicit diffusion
        a = RbcRhs()
        a.X, a.pres, a.cst, a.R = p(X), p(self.pres), p(self.st_const[st]), p(self.Rb)
        a.mdta, a.dta = p(self.st_mdta[st]), p(self.st_dta[st])
        a.fdiff = p(self.st_fdiff[st]) if self.diffusion else None
        a.tx, a.ty = p(self.tx), p(self.ty)
        a.M0, a.M1, a.N0, a.N1, a.batch = M0, M1, N0, N1, B
        self._keep.append(a)
        L.append(_Call("rhs_banded" + tag, lib.pypde_rbc_rhs_banded, ctypes.byref(a)))
        # 7  X_f = BF_x Z_f + R_f   (the forward transform along x of the
        #    convective term, onto the banded part)
        L.append(_Gemm("rhs" + tag, m=M0, n=M1, k=L0, batch=3 * B, A=p(self.BFx), lda=L0,
                       B=p(Z), ldb=M1, iB=_bidx(s1=sZ), D=p(self.Rb), ldd=M1, iD=_bidx(s1=sX),
                       beta=1.0, C=p(X), ldc=M1, iC=_bidx(s1=sX)))
        cands = {}
        for S in self._splits:
            kc = L0 // S
            part = _Gemm("rhs_part" + tag, m=M0, n=M1, k=kc, batch=3 * B * S, A=p(self.BFx),
                         lda=L0, iA=_bidx(s1=kc, m1=S), B=p(Z), ldb=M1,
                         iB=_bidx(s1=sZ, d1=S, s2=kc * M1, d2=1, m2=S),
                         C=p(self.rhs_work), ldc=M1, iC=_bidx(s1=sX))
            red = (ctypes.c_void_p(p(self.rhs_work)), S, ctypes.c_void_p(p(self.Rb)), None, 1, 0,
                   ctypes.c_void_p(p(X)), sX, 3 * B)
            cands[S] = _SplitRhs("rhs/%d" % S + tag, _GemmGroup([part]), red, S)
        self._rhs_pairs.append((L[-1], cands))
        # 9-10  Helmholtz ADI sweeps (register-resident recurrences)
        gx = _lines(p(X), p(X), M1, sX, M0, M1, 3 * B, 0)
        gy = _lines(p(X), p(X), M1, sX, M1, M0, 3 * B, 1)
        rx = {v: _rec(p(self.st_rhx[st]), M0, 4 * M0, v) for v in REC_VARIANTS}
        ry = {v: _rec(p(self.st_rhy[st]), M1, 4 * M1, v) for v in REC_VARIANTS}
        if self.exact_solves:
            for name, g, tab, M in (("helmholtz_x", gx, self.st_hx[st], M0),
                                    ("helmholtz_y", gy, self.st_hy[st], M1)):
                self._keep.append(g)
                t = tab.data_ptr()
                L.append(_Call(name + tag, lib.pypde_fdma_solve, ctypes.byref(g),
                               ctypes.c_void_p(t), ctypes.c_void_p(t + 8 * M),
                               ctypes.c_void_p(t + 16 * M), ctypes.c_void_p(t + 24 * M), 4 * M))
        else:
            L.append(_RecCall("helmholtz_x" + tag, gx, rx, _rec_default(3 * B * M1)))
            L.append(_RecCall("helmholtz_y" + tag, gy, ry, _rec_default(3 * B * M0)))
        # 11  W2_f = X*_f [C ; H_y C]_f^T  (f = U with C_y, V with Cd_y): the y factors
        #     of the divergence and of the Poisson right-hand side in one product
        NH = N1 + M1
        L.append(_Gemm("div_y" + tag, m=M0, n=NH, k=M1, batch=2 * B, A=p(X), lda=M1,
                       iA=_bidx(s1=3 * sX, d1=2, s2=sX, m2=2), B=p(self.HCy), ldb=M1, tb=1,
                       iB=_bidx(s1=NH * M1, m1=2), C=p(self.W2), ldc=NH, iC=_bidx(s1=M0 * NH)))
        # 12  div = [Cd | C]_x Wd  and  RP = (B [Cd | C])_x Wh  (neighbours: one launch)
        L.append(_Gemm("div_x" + tag, m=N0, n=N1, k=2 * M0, batch=B, A=p(self.Cx), lda=2 * M0,
                       B=p(self.W2), ldb=NH, iB=_bidx(s1=2 * M0 * NH), C=p(self.div), ldc=N1,
                       iC=_bidx(s1=N0 * N1)))
        L.append(_Gemm("poisson_rhs" + tag, m=M0, n=M1, k=2 * M0, batch=B, A=p(self.BCx),
                       lda=2 * M0, B=p(self.W2) + 8 * N1, ldb=NH, iB=_bidx(s1=2 * M0 * NH),
                       C=p(self.RP), ldc=M1, iC=_bidx(s1=sX)))
        # 15  per-line eigen-split solves along x
        gp = _lines(p(self.RP), p(self.RP), M1, sX, M0, M1, B, 0)
        rp = {}
        for v, tabs, layout in ((1, self.prec, 1), (2, self.precT, 2), (3, self.prec, 1),
                                (4, self.prec, 1)):
            r = Recurrence()
            r.LM, r.R, r.MU, r.NU = (tabs[i].data_ptr() for i in range(4))
            r.per_line, r.entry_stride, r.variant = layout, 0, v
            rp[v] = r
        if self.exact_solves:
            self._keep.append(gp)
            L.append(_Call("poisson_lines" + tag, lib.pypde_poisson_solve, ctypes.byref(gp),
                           ctypes.c_void_p(self.ptab.data_ptr()),
                           ctypes.c_void_p(self.psing.data_ptr()), 0))
        else:
            L.append(_RecCall("poisson_lines" + tag, gp, rp, _rec_default(B * M1)))
        # 16  P = RP Q_y^T with the gauge P[0,0] = 0 (GEMM epilogue), and
        # 17  W1 = P [S; T S; T D S/s]_y^T computed from RP directly (same launch)
        NW = N1 + 2 * M1
        L.append(_Gemm("poisson_qy" + tag, m=M0, n=M1, k=M1, batch=B, A=p(self.RP), lda=M1,
                       iA=_bidx(s1=sX), B=p(self.Qy), ldb=M1, tb=1, C=p(self.P), ldc=M1,
                       iC=_bidx(s1=sX), zero_first=1))
        L.append(_Gemm("corr_y" + tag, m=M0 - 1, n=NW, k=M1, batch=B, A=p(self.RP) + 8 * M1,
                       lda=M1, iA=_bidx(s1=sX), B=p(self.WQy), ldb=M1, tb=1,
                       C=p(self.W1) + 8 * NW, ldc=NW, iC=_bidx(s1=M0 * NW)))
        L.append(_Gemm("corr_y0" + tag, m=1, n=NW, k=M1, batch=B, A=p(self.RP), lda=M1,
                       iA=_bidx(s1=sX), B=p(self.WQy0), ldb=M1, tb=1, C=p(self.W1), ldc=NW,
                       iC=_bidx(s1=M0 * NW)))
        # 19-20  pres = (pres - nu div beta) + S_x W_0 / (dt a)   (one GEMM, fused epilogue)
        L.append(_Gemm("pres_p" + tag, m=N0, n=N1, k=M0, batch=B, A=p(self.Snx), lda=M0,
                       B=p(self.W1), ldb=NW, iB=_bidx(s1=M0 * NW), D=p(self.pres), ldd=N1,
                       iD=_bidx(s1=N0 * N1), beta=1.0, C=p(self.pres), ldc=N1,
                       iC=_bidx(s1=N0 * N1), alpha_vec=p(self.st_idta[st]),
                       E=p(self.div), lde=N1, iE=_bidx(s1=N0 * N1), gamma_vec=p(self.t_nu),
                       gamma=self.beta))
        # 21  velocity correction
        L.append(_Gemm("correct" + tag, m=M0, n=M1, k=M0, batch=2 * B, A=p(self.Acorr), lda=M0,
                       iA=_bidx(s1=M0 * M0, m1=2), B=p(self.W1) + 8 * N1, ldb=NW,
                       iB=_bidx(s1=M0 * NW, d1=2, s2=M1, m2=2), D=p(X), ldd=M1,
                       iD=_bidx(s1=3 * sX, d1=2, s2=sX, m2=2), beta=1.0, alpha=-1.0,
                       C=p(X), ldc=M1, iC=_bidx(s1=3 * sX, d1=2, s2=sX, m2=2)))

    def _yfold_launches(self, L, st, tag):
        """y-pass, products and y forward transform in the y-folded layout.  RK3
        stages alternate between two field buffers so that the c-term of stage
        st reads the velocities of stage st-1."""
        B, L0, L1, M1 = self.B, self.L0, self.L1, self.M1
        H, Mh = L1 // 2, M1 // 2
        p = lambda t: t.data_ptr()
        Y1, Pr, Z = self.Y1, self.Pr, self.Z
        Y2 = self.Y2 if st % 2 == 0 else self.Y2old
        old = None if st == 0 else (self.Y2old if st % 2 == 0 else self.Y2)
        ratio = float(self.kc[st]) / float(self.kb[st])
        yf = self.yfold
        sZ = self.sZ
        blk = L0 * H
        # 2a  value rows x {value, d/dy} x {even, odd}: entries e = 4 fi + q
        L.append(_Gemm("ypass_v" + tag, m=L0, n=H, k=Mh, batch=12 * B, A=p(Y1), lda=M1, ska=2,
                       iA=_bidx(s1=2 * L0 * M1, d1=4, s2=1, m2=2), B=p(yf["B"]), ldb=Mh, tb=1,
                       iB=_bidx(s1=H * Mh, m1=4), C=p(Y2), ldc=H,
                       iC=_bidx(s1=6 * blk, d1=4, s2=blk, m2=4)))
        # 2b  d/dx rows x value x {even, odd}: entries e = 2 fi + q
        L.append(_Gemm("ypass_x" + tag, m=L0, n=H, k=Mh, batch=6 * B, A=p(Y1) + 8 * L0 * M1,
                       lda=M1, ska=2, iA=_bidx(s1=2 * L0 * M1, d1=2, s2=1, m2=2), B=p(yf["B"]),
                       ldb=Mh, tb=1, iB=_bidx(s1=H * Mh, m1=2), C=p(Y2) + 8 * 4 * blk, ldc=H,
                       iC=_bidx(s1=6 * blk, d1=2, s2=blk, m2=2)))
        # 3  convective products at both mirrored columns (+ (c/b) times the
        #    products with the previous stage's velocities), folded for the
        #    forward transform
        L.append(_Call("products" + tag, lib.pypde_rbc_products_yfold, ctypes.c_void_p(p(Y2)),
                       ctypes.c_void_p(p(old) if old is not None else None),
                       ctypes.c_void_p(p(self.dTbcdz1)), ctypes.c_void_p(p(Pr)), L0, L1,
                       float(yf["sv"]), float(yf["sd"]), ratio, B))
        # 4  Z[0] = -dt b P BF_y^T: output columns q::2 from the sum (q = 0) or
        #    difference (q = 1) block when BF is mirror-even (sign > 0), else swapped
        fb = self.fold_bf
        if fb["sign"] > 0:
            a0, iA = p(Pr), _bidx(s1=2 * blk, d1=2, s2=blk, m2=2)
        else:
            a0, iA = p(Pr) + 8 * blk, _bidx(s1=2 * blk, d1=2, s2=-blk, m2=2)
        L.append(_Gemm("z_conv" + tag, m=L0, n=Mh, k=H, batch=6 * B, A=a0, lda=H, iA=iA,
                       B=p(fb["B"]), ldb=H, tb=1, iB=_bidx(s1=Mh * H, m1=2), C=p(Z), ldc=M1,
                       csn=2, iC=_bidx(s1=sZ, d1=2, s2=1, m2=2),
                       alpha_vec=p(self.st_mdtb[st]), alpha_div=6))

    def _plain_y_launches(self, L, tag):
        """y-pass, products and y forward transform in the natural layout (Euler)."""
        B, L0, L1, M1 = self.B, self.L0, self.L1, self.M1
        p = lambda t: t.data_ptr()
        Y1, Y2, Pr, Z = self.Y1, self.Y2, self.Pr, self.Z
        sZ = self.sZ
        # 2  y-pass: values and y-derivatives on the L1 grid
        g = _Gemm("ypass" + tag, m=2 * L0, n=2 * L1, k=M1, batch=3 * B, A=p(Y1), lda=M1,
                  iA=_bidx(s1=2 * L0 * M1), B=p(self.GGy), ldb=M1, tb=1,
                  C=p(Y2), ldc=2 * L1, iC=_bidx(s1=4 * L0 * L1))
        if self.fold_y is not None:
            fy, h = self.fold_y, M1 // 2
            g = self._with_alt(g, _Gemm(
                "ypass" + tag, m=2 * L0, n=L1, k=h, batch=3 * B, A=p(Y1), lda=M1, ska=2,
                iA=_bidx(s1=2 * L0 * M1), B=p(fy["E"]), ldb=h, tb=1,
                C=p(Y2), ldc=2 * L1, iC=_bidx(s1=4 * L0 * L1), fold=2, A1=p(Y1) + 8,
                B1=p(fy["O"]), lda1=M1, ldb1=h, k1=h, fold_len=L1, fold_sign=fy["sign"]))
        L.append(g)
        # 3  convective products
        L.append(_Call("products" + tag, lib.pypde_rbc_products, ctypes.c_void_p(p(Y2)),
                       ctypes.c_void_p(p(self.dTbcdz1)), ctypes.c_void_p(p(Pr)), L0, L1, B))
        # 4  Z[0] = -dt P BF_y^T
        g = _Gemm("z_conv" + tag, m=L0, n=M1, k=L1, batch=3 * B, A=p(Pr), lda=L1,
                  iA=_bidx(s1=L0 * L1), B=p(self.BFy), ldb=L1, tb=1, C=p(Z), ldc=M1,
                  iC=_bidx(s1=sZ), alpha_vec=p(self.st_mdtb[0]), alpha_div=3)
        if self.fold_bf is not None:
            fb, h = self.fold_bf, L1 // 2
            # entries e = 2 f + q: output columns q::2 of field-instance f
            g = self._with_alt(g, _Gemm(
                "z_conv" + tag, m=L0, n=M1 // 2, k=h, batch=6 * B, A=p(Pr), lda=L1,
                iA=_bidx(s1=L0 * L1, d1=2), B=p(fb["B"]), ldb=h, tb=1,
                iB=_bidx(s1=(M1 // 2) * h, m1=2), C=p(Z), ldc=M1, csn=2,
                iC=_bidx(s1=sZ, d1=2, s2=1, m2=2), alpha_vec=p(self.st_mdtb[0]), alpha_div=6,
                fold=3, fold_len=L1, mirror_sign=fb["sign"], mirror_div=1))
        L.append(g)


    # -----------------------------------------------------------------------
    def workspace_bytes(self):
        """Bytes of the workspace the step's launches touch: the K-split
        partials beyond the chosen number of slices are left out."""
        used = 8 * self.ws.numel()
        if self._splits:
            used = 8 * self.ws_offsets["rhs_work"]
            for l in self.launches:
                if isinstance(l, _SplitRhs):
                    used += 8 * 3 * self.B * l.S * self.M0 * self.M1
        return used

    def build_plan(self):
        """Record the step's launches (as tuned) into a step plan of the C-ABI
        (pypde_plan_*, regions X, pres, P and the workspace); pypde_rbc_step then
        runs one step in one call."""
        if self.plan is not None:
            return self.plan
        regions = [self.X, self.pres, self.P]
        bases = (ctypes.c_void_p * 4)(*[t.data_ptr() for t in regions], self.ws.data_ptr())
        sizes = (ctypes.c_longlong * 4)(*[8 * t.numel() for t in regions], self.workspace_bytes())
        self.plan = Plan(4, bases, sizes)
        for l in self.launches:
            parts = [l.group, l.reduce] if isinstance(l, _SplitRhs) else [l]
            for part in parts:
                self.plan.add(part)
        return self.plan

    def step_eager(self):
        """One step: the recorded step plan in one call (pypde_rbc_step)."""
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        plan = self.plan or self.build_plan()
        check(lib.pypde_rbc_step(plan.handle, ctypes.c_void_p(self.X.data_ptr()),
                                 ctypes.c_void_p(self.pres.data_ptr()),
                                 ctypes.c_void_p(self.P.data_ptr()),
                                 ctypes.c_void_p(self.ws.data_ptr()), 1, s), "pypde_rbc_step")

    def stage_of(self, launch):
        """The stage a launch of the step belongs to (names carry "[st]" when
        the scheme has several stages)."""
        if self.nstage == 1:
            return 0
        return int(launch.name.rsplit("[", 1)[1].rstrip("]"))

    def run_stage(self, st):
        """The launches of stage st alone, on the engine's buffers (the stages
        before it must have run on them)."""
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        for f in self.launches:
            if self.stage_of(f) == st:
                f(s)

    def step_launches(self):
        """One step as separate launches from Python (the plan's content)."""
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        for f in self.launches:
            f(s)

    def step_timed(self, records):
        """One eager step with a HIP event pair around every launch, on the
        current stream; appends (name, start, end, launch) to records."""
        stream = torch.cuda.current_stream()
        s = ctypes.c_void_p(stream.cuda_stream)
        for f in self.launches:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            f(s)
            e1.record(stream)
            records.append((f.name, e0, e1, f))

    @staticmethod
    def _with_alt(plain, folded):
        """The folded launch, remembering the plain one as the alternative."""
        folded.alt = plain
        return folded

    def tune(self):
        """Autotune the kernel variant of every GEMM launch of the step (the
        state is restored afterwards, so in-place launches can be timed)."""
        from . import autotune
        if not autotune.enabled():
            return
        saved = [t.clone() for t in (self.X, self.pres, self.P)]
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

        def launch_fn(name):
            return lambda d: check(lib.pypde_dgemm_batched(ctypes.byref(d), s), name)
        for launch in self.launches:
            if isinstance(launch, _RecCall):
                def run(v, launch=launch):
                    check(lib.pypde_recurrence_solve(ctypes.byref(launch.g),
                                                     ctypes.byref(launch.structs[v]), s),
                          launch.name)
                launch.set_variant(autotune.choose(launch.key(), run, REC_VARIANTS,
                                                   default=launch.variant))
        for launch in self.launches:
            if isinstance(launch, _Call) and launch.fn.__name__ == "pypde_rbc_rhs_banded":
                a = launch.args[0]._obj

                def run(v, launch=launch, a=a):
                    a.rows_per_thread = v
                    launch(s)
                a.rows_per_thread = autotune.choose(
                    ("rhs_banded_form", self.B, self.M0, self.M1, self.diffusion), run,
                    (1, 4, 8, 16), default=4)
        if self.group_launches:
            self._merge_products_rhs(s)
        for launch in self.launches:
            if isinstance(launch, _GemmGroup):
                self._tune_group(launch, s, launch_fn)
        for i, launch in enumerate(self.launches):
            if not isinstance(launch, _Gemm):
                continue
            autotune.tune_gemm_desc(launch.desc, launch_fn(launch.name), allow_alias=True)
            if launch.alt is not None:
                alt = launch.alt
                autotune.tune_gemm_desc(alt.desc, launch_fn(alt.name), allow_alias=True)
                forms = (launch, alt)
                pick = autotune.choose(
                    ("form",) + autotune.gemm_key(launch.desc) + autotune.gemm_key(alt.desc),
                    lambda c: launch_fn(forms[c].name)(forms[c].desc), (0, 1))
                if pick == 1:
                    alt.alt = launch
                    self.launches[i] = alt
        self._tune_rhs_split(s, launch_fn)
        torch.cuda.synchronize()
        for t, v in zip((self.X, self.pres, self.P), saved):
            t.copy_(v)
        torch.cuda.synchronize()

    def _merge_products_rhs(self, s):
        """Per stage, the y-folded products and the banded right-hand side
        (one output per thread) in one launch (pypde_rbc_products_rhs) where
        the tuner times that faster than the two launches, the right-hand side
        in the form tuned for it alone; the merged launch takes the products'
        place (the right-hand side reads only the stage-start state, which no
        launch in between changes)."""
        from . import autotune
        L = self.launches
        for st in range(self.nstage):
            tag = "" if self.nstage == 1 else "[%d]" % st
            names = {l.name: i for i, l in enumerate(L)}
            ip, ir = names.get("products" + tag), names.get("rhs_banded" + tag)
            if ip is None or ir is None or ip > ir:
                continue
            prod, rhs = L[ip], L[ir]
            if prod.fn.__name__ != "pypde_rbc_products_yfold":
                continue
            # the merged launch computes the right-hand side one output per thread
            a1 = RbcRhs.from_buffer_copy(rhs.args[0]._obj)
            a1.rows_per_thread = 1
            self._keep.append(a1)
            merged = _Call("products+rhs_banded" + tag, lib.pypde_rbc_products_rhs,
                           *(prod.args + (ctypes.byref(a1),) + rhs.args[1:]))
            forms = ((prod, rhs), (merged,))

            def run(c):
                for f in forms[c]:
                    f(s)
            if autotune.choose(("products_rhs", self.B, self.L0, self.L1, self.M0, self.M1,
                                self.diffusion, st > 0), run, (0, 1), default=1) == 1:
                L[ip] = merged
                del L[ir]

    def _tune_group(self, launch, s, launch_fn):
        """Variant of a grouped launch, or separate launches when faster."""
        from . import autotune
        for m in launch.members:
            autotune.tune_gemm_desc(m.desc, launch_fn(m.name), allow_alias=True)
        own = [m.desc.variant for m in launch.members]

        def run_group(v, g=launch):
            g.set_variant(v)
            check(lib.pypde_dgemm_grouped(ctypes.byref(g.arr), len(g.members), s), g.name)
        keys = tuple(autotune.gemm_key(m.desc) for m in launch.members)
        best = autotune.choose(("group",) + keys, run_group, autotune.GEMM_CANDIDATES,
                               default=own[0])

        def run_mode(c, g=launch):
            if c == 0:
                g.set_variant(best)
                check(lib.pypde_dgemm_grouped(ctypes.byref(g.arr), len(g.members), s), g.name)
            else:
                for m, v in zip(g.members, own):
                    m.desc.variant = v
                    m(s)
        separate = autotune.choose(("group_mode",) + keys, run_mode, (0, 1), default=0)
        if separate:
            for m, v in zip(launch.members, own):
                m.desc.variant = v
            launch.split = True
        else:
            launch.set_variant(best)
            launch.split = False

    def _tune_rhs_split(self, s, launch_fn):
        """Launches 7-8 as they are (0) or split along K into S slices."""
        from . import autotune
        for rhs, cands in self._rhs_pairs:
            if not cands:
                continue
            for c in cands.values():
                self._tune_group(c.group, s, launch_fn)

            def run(S, rhs=rhs, cands=cands):
                if S == 0:
                    rhs(s)
                else:
                    cands[S](s)
            key = ("rhs_split_banded", self.B, self.M0, self.M1, self.L0,
                   autotune.gemm_key(rhs.desc))
            pick = autotune.choose(key, run, (0,) + tuple(cands), default=0)
            if pick:
                i = self.launches.index(rhs)
                self.launches[i] = cands[pick]

    def variants(self):
        """{launch name: kernel variant} chosen for this step (GEMM variants;
        grouped launches: their variant or "split"; recurrences: 1 block form,
        2 lane-chunk form, 3 block form with 32 lines per block)."""
        out = {l.name: l.desc.variant for l in self.launches if isinstance(l, _Gemm)}
        out.update({l.name: ("split" if l.split else l.variant) for l in self.launches
                    if isinstance(l, _GemmGroup)})
        out.update({l.name: l.variant for l in self.launches if isinstance(l, _RecCall)})
        return out

    def capture(self, nsteps=1):
        """Capture `nsteps` consecutive steps into one HIP graph (the state is
        left unchanged: the warm-up step before capture is undone)."""
        saved = [t.clone() for t in (self.X, self.pres, self.P)]
        self.step_eager()  # warm up outside capture
        torch.cuda.synchronize()
        for t, v in zip((self.X, self.pres, self.P), saved):
            t.copy_(v)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for _ in range(nsteps):
                self.step_eager()
        # one replay now, so that no later replay is the graph's first, then
        # the state is put back
        g.replay()
        torch.cuda.synchronize()
        for t, v in zip((self.X, self.pres, self.P), saved):
            t.copy_(v)
        torch.cuda.synchronize()
        self.graph = g
        self.graph_steps = nsteps
        self.graphs[nsteps] = g

    def step(self, n=1):
        """Advance n steps: replays of the longest captured graphs that fit (a
        graph of several steps costs one launch instead of one per step), the
        rest as launches.  Replays run exactly the captured launches, so every
        split of n gives the same state bit for bit."""
        for length in sorted(self.graphs, reverse=True):
            g = self.graphs[length]
            for _ in range(n // length):
                g.replay()
            n %= length
        for _ in range(n):
            self.step_eager()

    def gemm_flops(self):
        return sum(l.flops for l in self.launches if isinstance(l, _Gemm))


def transfer_function(TL, TM, TR, x, k=0.01):
    """Sidewall temperature profile of the "zero" case (rbc2d.py:437-446)."""
    arr = np.zeros(x.shape)
    L = x[-1] - x[0]
    for i in range(x.size):
        xs = x[i] * 2.0 / L
        if xs < 0:
            arr[i] = -k * xs / (k + xs + 1) * (TL - TM) + TM
        else:
            arr[i] = k * xs / (k - xs + 1) * (TR - TM) + TM
    return arr
