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
import numpy as np
import torch

from .engine import RBCBatch


def default_dt(ra):
    """Per-instance time step rule for batched runs: dt = 0.01 sqrt(1e5 / Ra),
    clipped to [1e-3, 1e-2] (0.01 at Ra = 1e5, 1e-3 at Ra = 1e7 and above)."""
    return float(np.clip(0.01 * np.sqrt(1e5 / ra), 1e-3, 1e-2))


class HeteroBatch:
    """specs: sequence of dicts with keys shape=(N0, N1), ra, pr, dt."""

    def __init__(self, specs, case="rbc", dealias=True, aspect=1.0):
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
            self.classes.append(RBCBatch(shape, params, case=case, dealias=dealias,
                                         aspect=aspect))
            for slot, i in enumerate(idx):
                self.where[i] = (c, slot)
        self.graph = None
        self.graph_steps = 0
        self._streams = None

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

    # ---- stepping ---------------------------------------------------------
    def step_eager(self):
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
        self._streams = [torch.cuda.Stream() for _ in self.classes]
        g = torch.cuda.CUDAGraph()
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
