"""Heterogeneous batches: many independent RBC instances with different
resolutions, Rayleigh / Prandtl numbers and time steps, advanced together by
one HIP-graph launch per step (or per `nsteps` steps).

Instances are grouped by shape into size classes; each class is an
`engine.RBCBatch` (one launch per operator for all its instances, GEMM and
banded schedules autotuned for that shape).  The captured graph forks one
stream branch per size class from the capture stream and joins them at the
end of every step, so classes of small fields overlap on the device instead of
running one after another.  No reference counterpart: the reference steps one
NavierStokes object at a time (navier/rbc2d.py:396-434).
"""
import ctypes

import numpy as np
import torch

from ._lib import Lines, MultiInfo, RbcRhs, Recurrence, check, lib
from .engine import _GROUPS, RBCBatch, _Call, _Gemm, _RecCall

# the LDS-tiled GEMM variants a multi launch can run (one for all members)
MULTI_GEMM_VARIANTS = (1, 2, 3, 5, 6, 7, 24, 25, 26, 27, 28, 33, 34, 35, 36, 37, 38)
_KIND = {"gemm": 1, "recurrence": 2, "products": 3, "rhs": 4}


class _Multi:
    """One multi launch of the C-ABI (pypde_*_multi): the members' arguments
    packed on the host, copied once into a device table, launched in one go."""

    def __init__(self, name, kind, pack):
        self.name, self.kind = name, kind
        self.info = MultiInfo()
        nbytes = lib.pypde_multi_table_bytes(_KIND[kind])
        self.host = (ctypes.c_ubyte * nbytes)()
        self.pack = pack
        self.table = None
        self.repack()

    def repack(self, *args):
        check(self.pack(self.host, len(self.host), ctypes.byref(self.info), *args), self.name)
        host = torch.frombuffer(bytearray(self.host), dtype=torch.uint8)
        dev = torch.empty(host.numel(), dtype=torch.uint8, device="cuda")
        dev.copy_(host)
        torch.cuda.synchronize()
        self.table = dev  # the launch reads it: kept alive with the launch

    def __call__(self, stream):
        fn = {"gemm": lib.pypde_dgemm_multi, "recurrence": lib.pypde_recurrence_multi,
              "products": lib.pypde_products_multi, "rhs": lib.pypde_rhs_multi}[self.kind]
        check(fn(ctypes.c_void_p(self.table.data_ptr()), ctypes.byref(self.info), stream),
              self.name)


def default_dt(ra):
    """Per-instance time step rule for batched runs: dt = 0.01 sqrt(1e5 / Ra),
    clipped to [1e-3, 1e-2] (0.01 at Ra = 1e5, 1e-3 at Ra = 1e7 and above)."""
    return float(np.clip(0.01 * np.sqrt(1e5 / ra), 1e-3, 1e-2))


class HeteroBatch:
    """specs: sequence of dicts with keys shape=(N0, N1), ra, pr, dt.

    persistent=True: every operator of the step runs once for all size classes
    (the multi launches of the C-ABI, one per stage: about 13 per Euler step
    whatever the number of classes); False: one engine step per class on its
    own stream branch of the captured graph.

    groups (persistent form): the size classes are split into this many groups
    of about equal work, each with its own chain of multi launches; the captured
    graph runs the chains on parallel branches (13 x groups nodes per step), so
    that one group's latency-bound line solves overlap another's products."""

    def __init__(self, specs, case="rbc", dealias=True, aspect=1.0, persistent=False, groups=1):
        self.specs = [dict(s) for s in specs]
        if not self.specs:
            raise ValueError("HeteroBatch needs at least one instance")
        self.case = case
        by_shape = {}
        for i, s in enumerate(self.specs):
            shape = tuple(int(n) for n in s["shape"])
            if len(shape) != 2:
                raise ValueError("instance %d: shape must be (N0, N1)" % i)
            by_shape.setdefault(shape, []).append(i)
        self.shapes = sorted(by_shape)
        self.classes = []
        self.where = [None] * len(self.specs)  # instance -> (class, slot)
        for c, shape in enumerate(self.shapes):
            idx = by_shape[shape]
            params = [dict(ra=self.specs[i]["ra"], pr=self.specs[i].get("pr", 1.0),
                           dt=self.specs[i]["dt"]) for i in idx]
            # the persistent form takes each class's launches one by one (no
            # grouped or K-split launches inside a class)
            self.classes.append(RBCBatch(shape, params, case=case, dealias=dealias,
                                         aspect=aspect, group_launches=not persistent))
            for slot, i in enumerate(idx):
                self.where[i] = (c, slot)
        self.graph = None
        self.graph_steps = 0
        self._streams = None
        self.persistent = bool(persistent)
        self.groups = None
        self.multi = None
        if self.persistent:
            self.groups = self._partition(max(1, min(int(groups), len(self.classes))))
            self.group_multi = [self._build_multi(g) for g in self.groups]
            self.multi = [m for chain in self.group_multi for m in chain]

    def __len__(self):
        return len(self.specs)

    # ---- state ------------------------------------------------------------
    def temperature_field(self, c):
        """The Field of the temperature coefficients of size class c."""
        from .bases import Base
        from .field import Field
        N0, N1 = self.shapes[c]
        return Field([Base(N0, "CN" if self.case == "rbc" else "CD"), Base(N1, "CD")])

    def seed(self, seeds, amplitude=1e-3):
        """T of instance i = forward(uniform noise of the given amplitude on the
        N0 x N1 grid, default_rng(seeds[i])); U = V = P = pres = 0."""
        if len(seeds) != len(self):
            raise ValueError("one seed per instance")
        fields = [self.temperature_field(c) for c in range(len(self.classes))]
        for eng in self.classes:
            eng.X.zero_()
            eng.pres.zero_()
            eng.P.zero_()
        for i, sd in enumerate(seeds):
            c, slot = self.where[i]
            rng = np.random.default_rng(int(sd))
            noise = rng.uniform(-amplitude, amplitude, size=self.shapes[c])
            self.classes[c].X[slot, 2].copy_(fields[c].forward(noise))

    def state(self, i):
        """Views of instance i: dict(U, V, T, pres) (Galerkin / Chebyshev coefficients)."""
        c, slot = self.where[i]
        e = self.classes[c]
        return dict(U=e.X[slot, 0], V=e.X[slot, 1], T=e.X[slot, 2], pres=e.pres[slot])

    def finite(self):
        return all(bool(torch.isfinite(e.X).all().item()) and bool(torch.isfinite(e.pres).all().item())
                   for e in self.classes)

    def nusselt(self):
        """Plate Nusselt number of every instance (instance order)."""
        from .adjoint import NusseltFunctional
        out = np.empty(len(self))
        for c, e in enumerate(self.classes):
            vals = NusseltFunctional(e).value(e.X[:, 2]).detach().cpu().numpy()
            for i, (cc, slot) in enumerate(self.where):
                if cc == c:
                    out[i] = vals[slot]
        return out

    # ---- one launch per stage for every class -----------------------------
    def _partition(self, n):
        """Class indices in n groups of about equal work (largest first onto the
        lightest group); the work of a class is its instances times its
        dealiased grid times its Galerkin extent, the shape of the GEMM work."""
        work = [e.B * e.L0 * e.L1 * (e.M0 + e.M1) for e in self.classes]
        groups, load = [[] for _ in range(n)], [0] * n
        for c in sorted(range(len(work)), key=lambda c: -work[c]):
            g = load.index(min(load))
            groups[g].append(c)
            load[g] += work[c]
        return [sorted(g) for g in groups]

    def _build_multi(self, cls=None):
        """The step as multi launches over the classes `cls` (all by default):
        position p of every class's launch list (the same operator of the step)
        becomes one launch, and neighbouring GEMM positions that the engine may
        group share it."""
        cls = list(range(len(self.classes))) if cls is None else list(cls)
        lists = [self.classes[c].launches for c in cls]
        names = [l.name for l in lists[0]]
        for L in lists[1:]:
            if [l.name for l in L] != names:
                raise RuntimeError("size classes issue different launch sequences")
        mergeable = set()
        for grp in _GROUPS:
            for a, b in zip(grp, grp[1:]):
                mergeable.add((a, b))
        base = lambda name: name.split("[", 1)[0]   # without the RK3 stage tag
        saved = [[t.clone() for t in (self.classes[c].X, self.classes[c].pres, self.classes[c].P)]
                 for c in cls]
        out, p = [], 0
        while p < len(names):
            first = lists[0][p]
            if isinstance(first, _Gemm):
                q = p + 1
                while (q < len(names) and isinstance(lists[0][q], _Gemm)
                       and (base(names[q - 1]), base(names[q])) in mergeable
                       and (self._plain(lists[0][q]).desc.transA,
                            self._plain(lists[0][q]).desc.transB) ==
                       (self._plain(first).desc.transA, self._plain(first).desc.transB)):
                    q += 1
                descs = [self._plain(L[k]).desc for k in range(p, q) for L in lists]
                folded = [self._folded(L[p]) for L in lists] if q == p + 1 else [None]
                if all(f is not None for f in folded) and \
                        len({f.desc.fold for f in folded}) == 1:
                    out.append(self._multi_gemm(names[p], descs, [f.desc for f in folded]))
                else:
                    out.append(self._multi_gemm("+".join(names[p:q]), descs))
                p = q
                continue
            members = [L[p] for L in lists]
            if isinstance(first, _RecCall):
                out.append(self._multi_rec(names[p], members))
            elif isinstance(first, _Call) and first.fn.__name__ == "pypde_rbc_products_yfold":
                out.append(self._multi_products(names[p], members))
            elif isinstance(first, _Call) and first.fn.__name__ == "pypde_rbc_rhs_banded":
                out.append(self._multi_rhs(names[p], members))
            else:
                raise RuntimeError("launch %s has no multi form" % names[p])
            p += 1
        torch.cuda.synchronize()
        for c, sv in zip(cls, saved):
            e = self.classes[c]
            for t, v in zip((e.X, e.pres, e.P), sv):
                t.copy_(v)
        torch.cuda.synchronize()
        return out

    @staticmethod
    def _plain(launch):
        """The unfolded form of a GEMM launch (multi launches run plain tiles)."""
        if launch.desc.fold != 0:
            if launch.alt is None:
                raise RuntimeError("%s has no unfolded form" % launch.name)
            return launch.alt
        return launch

    @staticmethod
    def _folded(launch):
        """The parity-folded form of a GEMM launch, if it has one."""
        if launch.desc.fold != 0:
            return launch
        if launch.alt is not None and launch.alt.desc.fold != 0:
            return launch.alt
        return None

    def _multi_gemm(self, name, descs, folded=None):
        """One multi launch of the plain descriptors (variants MULTI_GEMM_VARIANTS)
        or, when given, of their folded forms (variants 1-5: half the MFMA work
        of a transform), whichever form and variant the tuner times fastest."""
        from . import autotune
        from ._lib import GemmDesc
        arrs = {"plain": (GemmDesc * len(descs))(*descs)}
        if folded is not None:
            arrs["fold"] = (GemmDesc * len(folded))(*folded)
        state = {"form": "plain", "v": 2}

        def pack(host, nb, info):
            arr = arrs[state["form"]]
            return lib.pypde_gemm_multi_pack(ctypes.byref(arr), len(arr), state["v"], host, nb,
                                             info)
        m = _Multi(name, "gemm", pack)
        m._arr = arrs
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        tables = {}

        def run(c):
            if c not in tables:
                state["form"], state["v"] = c
                m.repack()
                tables[c] = (m.table, MultiInfo.from_buffer_copy(m.info))
            t, info = tables[c]
            check(lib.pypde_dgemm_multi(ctypes.c_void_p(t.data_ptr()), ctypes.byref(info), stream),
                  name)
        cands = [("plain", v) for v in MULTI_GEMM_VARIANTS]
        key = ("multi_gemm",) + tuple(autotune.gemm_key(d) for d in descs)
        if folded is not None:
            cands += [("fold", v) for v in (1, 2, 3, 4, 5)]
            key = ("multi_gemm_or_fold",) + key[1:] + tuple(autotune.gemm_key(d) for d in folded)
        best = autotune.choose(key, run, cands, default=("plain", 2))
        state["form"], state["v"] = best
        m.repack()
        m.variant = best[1]
        m.form = best[0]
        return m

    def _multi_rec(self, name, members):
        n = len(members)
        gs = (Lines * n)(*[l.g for l in members])
        rs = (Recurrence * n)(*[l.structs[2] for l in members])
        m = _Multi(name, "recurrence", lambda host, nb, info: lib.pypde_recurrence_multi_pack(
            ctypes.byref(gs), ctypes.byref(rs), n, host, nb, info))
        m._keep = (gs, rs)
        return m

    def _multi_products(self, name, members):
        n = len(members)
        cols = list(zip(*[l.args for l in members]))   # Y, Yold, dtbc, P, L0, L1, sv, sd, ratio, B
        vp = ctypes.c_void_p
        arrs = [(vp * n)(*[a.value if a is not None else None for a in cols[i]]) for i in range(4)]
        ints = [(ctypes.c_int * n)(*cols[i]) for i in (4, 5)]
        dbls = [(ctypes.c_double * n)(*cols[i]) for i in (6, 7, 8)]
        bat = (ctypes.c_int * n)(*cols[9])
        m = _Multi(name, "products", lambda host, nb, info: lib.pypde_products_multi_pack(
            *arrs, *ints, *dbls, bat, n, host, nb, info))
        m._keep = (arrs, ints, dbls, bat)
        return m

    def _multi_rhs(self, name, members):
        """The banded right-hand sides of every class in one launch, in the
        kernel form (one output or a strip of 4 / 8 rows per thread, or LDS
        tiles) the tuner times fastest for the whole set."""
        from . import autotune
        n = len(members)
        args = (RbcRhs * n)(*[RbcRhs.from_buffer_copy(l.args[0]._obj) for l in members])

        def pack(host, nb, info):
            return lib.pypde_rhs_multi_pack(ctypes.byref(args), n, host, nb, info)

        def set_form(rs):
            for a in args:
                a.rows_per_thread = rs
        set_form(1)
        m = _Multi(name, "rhs", pack)
        m._keep = args
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        tables = {}

        def run(rs):
            if rs not in tables:
                set_form(rs)
                m.repack()
                tables[rs] = (m.table, MultiInfo.from_buffer_copy(m.info))
            t, info = tables[rs]
            check(lib.pypde_rhs_multi(ctypes.c_void_p(t.data_ptr()), ctypes.byref(info), stream),
                  name)
        key = ("multi_rhs",) + tuple((a.M0, a.M1, a.batch, a.fdiff is not None) for a in args)
        best = autotune.choose(key, run, (1, 4, 8, 16), default=1)
        set_form(best)
        m.repack()
        return m

    def build_plan(self):
        """The persistent step as a step plan of the C-ABI (its multi launches,
        run in place on the classes' buffers by pypde_rbc_step_batched)."""
        from .engine import Plan
        if not self.persistent:
            raise RuntimeError("only the persistent form has a step plan")
        if getattr(self, "plan", None) is None:
            plan = Plan(0, None, None)
            for m in self.multi:
                check(lib.pypde_plan_add_multi(plan.handle, ctypes.c_void_p(m.table.data_ptr()),
                                               ctypes.byref(m.info)), m.name)
            self.plan = plan
        return self.plan

    # ---- stepping ---------------------------------------------------------
    def step_eager(self):
        if self.persistent:
            s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
            check(lib.pypde_rbc_step_batched(self.build_plan().handle, 1, s),
                  "pypde_rbc_step_batched")
            return
        for e in self.classes:
            e.step_eager()

    def capture(self, nsteps=1):
        """Capture `nsteps` steps of every size class into one HIP graph (the
        state is left unchanged: the warm-up step before capture is undone)."""
        saved = [[t.clone() for t in (e.X, e.pres, e.P)] for e in self.classes]
        self.step_eager()
        torch.cuda.synchronize()
        for e, sv in zip(self.classes, saved):
            for t, v in zip((e.X, e.pres, e.P), sv):
                t.copy_(v)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        if self.persistent and len(self.group_multi) == 1:
            with torch.cuda.graph(g):
                for _ in range(nsteps):
                    self.step_eager()
            self.graph = g
            self.graph_steps = nsteps
            return
        if self.persistent:
            # one branch per group, each its chain of multi launches
            self._streams = [torch.cuda.Stream() for _ in self.group_multi]
            with torch.cuda.graph(g):
                cap = torch.cuda.current_stream()
                for _ in range(nsteps):
                    for s, chain in zip(self._streams, self.group_multi):
                        s.wait_stream(cap)
                        with torch.cuda.stream(s):
                            sp = ctypes.c_void_p(s.cuda_stream)
                            for m in chain:
                                m(sp)
                    for s in self._streams:
                        cap.wait_stream(s)
            self.graph = g
            self.graph_steps = nsteps
            return
        self._streams = [torch.cuda.Stream() for _ in self.classes]
        with torch.cuda.graph(g):
            cap = torch.cuda.current_stream()
            for _ in range(nsteps):
                for s, e in zip(self._streams, self.classes):
                    s.wait_stream(cap)
                    with torch.cuda.stream(s):
                        e.step_eager()
                for s in self._streams:
                    cap.wait_stream(s)
        self.graph = g
        self.graph_steps = nsteps

    def step(self, n=1):
        """Advance every instance by n steps (graph replays when captured)."""
        if self.graph is None:
            for _ in range(n):
                self.step_eager()
            return
        if n % self.graph_steps:
            raise ValueError("n must be a multiple of the captured step count %d" % self.graph_steps)
        for _ in range(n // self.graph_steps):
            self.graph.replay()
