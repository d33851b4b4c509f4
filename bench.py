#!/usr/bin/env python
"""Benchmark: RBC2D FP64 time steps per second at 256x256 on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra]
    python bench.py --micro              # microbench only (one JSON line)
    python bench.py --adjoint            # adjoint gradient at 256x256 only (one JSON line)
    python bench.py --copy-calibration   # 256 MiB device copies, for scaling PMC byte counters
    python bench.py --tune-db PATH       # tune every workload afresh, write the decisions
    python bench.py --gemm-k-sweep       # y-pass-shaped GEMM against K (fixed cost per launch)

Workload (value): the plain Euler IMEX step of navier.rbc2d.NavierStokes at
256x256, rbc case, Ra=1e8, Pr=1, dt=1e-3, 3/2 dealiasing, seeded
conductive-plus-noise initial state; W untimed steps, then K steps timed
between device synchronisations, replayed from HIP graphs of ten steps
(NavierStokes.update_many).

Also reported on the same line:
  roofline       the dominant launch of the timed step (largest time per step in an
                 eager pass of the same K steps with HIP events around each launch),
                 timed as 20 back-to-back launches replayed from a HIP graph; for the
                 K-split right-hand side, its GEMM kernel (partial products and the
                 buoyancy product in one launch)
  batched        64 instances of 256x256 (Ra log-uniform in [1e4, 1e7]) in one
                 batched step: aggregate instance-steps/s and the ratio to the plain path
  variants       the same 256x256 step with RK3 (three stages) and with beta = 0.5
  hetero         256 instances of mixed sizes (N per axis in {64, 96, 128, 192}) and
                 Ra, one graph launch per step, 500 steps: aggregate instance-steps/s and
                 the ratio to looping the plain single-instance step over them; the
                 persistent form (one launch per stage for all 16 size classes), and
                 the stream-branch form (one branch per class) as hetero_branches
  microbench     512 fields of 256x256: transforms, ChebDirichlet stencil maps,
                 Helmholtz right-hand side and four-diagonal solves, per-kernel HBM fractions
  adjoint        gradient of the time-averaged plate Nusselt number with respect
                 to the initial temperature (checkpointed reverse sweep): 200 steps
                 at 256x256 and 2000 steps at 128x128, Ra=5e4, dt=5e-3, checkpoints
                 every 100 steps; reported as trajectory steps/s of the whole gradient
  steady         time to a steady state by Newton-Krylov on the fused step (20x16 and
                 64x64 at Ra=1e4, 128x128 at Ra=5e4; 300 steps from a roll, then
                 max|F| <= 1e-10 with F over 0.5 time units)
  adjoint_descent  the reference's adjoint-descent update (navier/rbc2d_adj.py) at 256x256
  cpu_baseline   the oracle (CPU restatement of the reference) on the host, one
                 thread, on a bounded sample of the same 256x256 workload
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def seeded_noise(shape, seed=0, amplitude=1e-3):
    rng = np.random.default_rng(seed)
    return rng.uniform(-amplitude, amplitude, size=shape)


# HBM bytes per launch of each launch of the 256x256 step (kernel choices of
# pypde_amd/tune_db.json), from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes
# over `python3 bench.py --steps 20 --warmup 2 --no-cpu --no-extra
# --all-launches` (the mean over 20 graph-replayed launches of each), each
# counter divided by what it reads for a device copy of known size
# (`python3 bench.py --copy-calibration`); profiles/hbm_traffic_pmc.md gives the
# commands, the calibration and the per-launch table.
TRAFFIC_BYTES = {
    "xpass": 12615994,
    "ypass_v+ypass_x": 23615960,
    "products+rhs_banded": 27793527,
    "z_conv": 10696155,
    "rhs": 13369201,
    "helmholtz_x": 3658500,
    "helmholtz_y": 3514573,
    "div_y": 11533202,
    "div_x+poisson_rhs": 12059385,
    "poisson_lines": 3499907,
    "poisson_qy+corr_y+corr_y0": 15145603,
    "pres_p+correct": 10773292,
}


def launch_graph_time(eng, launch, reps=20):
    """Seconds per launch of one launch of the step, replayed back to back from a
    HIP graph (as in the captured step); the engine state is restored."""
    import ctypes
    import torch
    saved = [t.clone() for t in (eng.X, eng.pres, eng.P)]
    launch(ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        for _ in range(reps):
            launch(st)
    g.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    g.replay()
    e1.record()
    e1.synchronize()
    for t, v in zip((eng.X, eng.pres, eng.P), saved):
        t.copy_(v)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) * 1e-3 / reps


def copy_calibration(mib=256, reps=5):
    import torch
    n = mib * 1024 * 1024 // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read": 8 * n, "bytes_written": 8 * n,
                                           "launches": reps, "kernel": "torch.mul(x, 1.0, out=y)"}}))


def launch_cost(name, launch, eng):
    """(bound, algorithmic amount per launch) of one launch of the fused step."""
    from pypde_amd.engine import _Gemm, _GemmGroup
    if isinstance(launch, (_Gemm, _GemmGroup)):
        return "mfma", launch.flops
    B, N0, N1, M0, M1, L0, L1 = eng.B, eng.N0, eng.N1, eng.M0, eng.M1, eng.L0, eng.L1
    prod = (3 * 3 * L0 * L1 + 3 * L0 * L1) * B + L0 * L1
    rhs = (3 * M0 * M1 + N0 * N1 + 3 * M0 * M1 + 3 * M0 * M1) * B
    el = {
        # read 9 blocks + the lifted term, write 3 products
        "products": prod,
        "products+rhs_banded": prod + rhs,
        "helmholtz_x": 2 * 3 * M0 * M1 * B,
        "helmholtz_y": 2 * 3 * M0 * M1 * B,
        # in/out plus the per-line factor tables, counted once per launch
        "poisson_lines": 2 * M0 * M1 * B + 4 * M0 * M1,
        # read X, pres and the constant terms, write R
        "rhs_banded": rhs,
    }.get(name, 0)
    return "hbm", 8.0 * el


def time_plain(args):
    import torch
    from pypde_amd.navier import NavierStokes
    N = 256
    ns = NavierStokes(case="rbc", shape=(N, N), ra=1e8, pr=1.0, dt=1e-3, tsave=None,
                      dealias=True, integrator="eu", beta=1.0, aspect=1.0)
    ns.set_state(T=ns.T.forward(seeded_noise((N, N), seed=0)))
    # every graph the timed region replays is captured (and replayed once, the
    # state restored) here, whatever W is
    ns.prepare_graphs()
    ns.update_many(args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ns.update_many(args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sec = t1 - t0
    # kernel pass: the same K steps eagerly with events around each launch
    eng = ns.engine
    recs = []
    for _ in range(args.steps):
        eng.step_timed(recs)
    torch.cuda.synchronize()
    per = {}
    for name, e0, e1, launch in recs:
        per.setdefault(name, [0.0, 0, launch])
        per[name][0] += e0.elapsed_time(e1) * 1e-3
        per[name][1] += 1
    # the dominant launch: the longest when each launch is replayed back to back
    # from a HIP graph (as in the timed steps); the eager event times above also
    # carry the host's launch gaps
    from pypde_amd.engine import _SplitRhs
    graph_s = {}
    for l in eng.launches:
        for part in ([l.group, l.reduce] if isinstance(l, _SplitRhs) else [l]):
            graph_s[part.name] = (launch_graph_time(eng, part), part)
    dom = max(graph_s, key=lambda k: graph_s[k][0])
    avg, launch = graph_s[dom]
    bound, amount = launch_cost(dom, launch, eng)
    if bound == "mfma":
        achieved, peak, unit = amount / avg / 1e12, FP64_PEAK_TFLOPS, "TFLOP/s"
    else:
        achieved, peak, unit = amount / avg / 1e9, HBM_PEAK_GBS, "GB/s"
    kernel_table = {k: dict(avg_us=round(v[0] / v[1] * 1e6, 3), calls=v[1]) for k, v in per.items()}
    for k, (t, _) in graph_s.items():
        kernel_table.setdefault(k, {})["graph_us"] = round(t * 1e6, 3)
    nu = ns.eval_Nu()
    finite = bool(np.isfinite(nu[0]) and np.isfinite(nu[1]))
    return dict(sec=sec, steps_per_s=args.steps / sec, dominant=dom, avg_s=avg, bound=bound,
                amount=amount, achieved=achieved, peak=peak, unit=unit, kernels=kernel_table,
                nu=nu, finite=finite, eager_sec=sum(v[0] for v in per.values()), engine=eng)


def time_batched(args, nb=64, steps=None):
    import torch
    from pypde_amd.engine import RBCBatch
    steps = steps or max(10, args.steps // 4)
    N = 256
    rng = np.random.default_rng(2024)
    ras = 10.0 ** rng.uniform(4, 7, size=nb)
    dts = np.clip(0.01 * np.sqrt(1e5 / ras), 1e-3, 1e-2)
    params = [dict(ra=float(r), pr=1.0, dt=float(d)) for r, d in zip(ras, dts)]
    eng = RBCBatch((N, N), params, case="rbc")
    from pypde_amd.field import Field
    from pypde_amd.bases import Base
    Tf = Field([Base(N, "CN"), Base(N, "CD")])
    for b in range(nb):
        eng.X[b, 2].copy_(Tf.forward(seeded_noise((N, N), seed=1000 + b)))
    eng.capture(1)
    eng.step(3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.step(steps)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    finite = bool(torch.isfinite(eng.X).all().item())
    return dict(instances=nb, steps=steps, sec=sec, inst_steps_per_s=nb * steps / sec,
                finite=finite)


def c4_specs(n=256, seed=7):
    """The mixed-size batch: N per axis uniform from {64, 96, 128, 192}, Ra
    log-uniform in [1e4, 1e7], Pr = 1, dt = clip(0.01 sqrt(1e5/Ra), 1e-3, 1e-2);
    instance i is seeded with noise 1e-3 from seed 1000 + i."""
    from pypde_amd.batch import default_dt
    rng = np.random.default_rng(seed)
    sizes = np.array([64, 96, 128, 192])
    nx = rng.choice(sizes, size=n)
    ny = rng.choice(sizes, size=n)
    ras = 10.0 ** rng.uniform(4, 7, size=n)
    return [dict(shape=(int(a), int(b)), ra=float(r), pr=1.0, dt=default_dt(r))
            for a, b, r in zip(nx, ny, ras)]


def time_hetero(steps=500, warm=5, n=256, seed=7, persistent=True, groups=1):
    """256 independent instances, N drawn per axis from {64, 96, 128, 192},
    Ra log-uniform in [1e4, 1e7], Pr = 1, dt = clip(0.01 sqrt(1e5/Ra), 1e-3, 1e-2),
    instance i seeded with noise 1e-3 from seed 1000 + i; one graph launch per
    step for all size classes.  Compared with looping the plain single-instance
    graph-replayed step over the same instances."""
    import torch
    from pypde_amd.batch import HeteroBatch
    from pypde_amd.engine import RBCBatch
    specs = c4_specs(n, seed)
    t0 = time.perf_counter()
    hb = HeteroBatch(specs, case="rbc", persistent=persistent, groups=groups)
    setup = time.perf_counter() - t0
    hb.seed([1000 + i for i in range(n)], amplitude=1e-3)
    hb.capture(1)
    hb.step(warm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hb.step(steps)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    finite = hb.finite()
    # plain path: one graph-replayed single instance per shape, timed per step
    per_shape = {}
    for shape in hb.shapes:
        e = RBCBatch(shape, [dict(ra=1e5, pr=1.0, dt=0.01)], case="rbc")
        e.capture(1)
        e.step(5)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        e.step(100)
        torch.cuda.synchronize()
        per_shape[shape] = (time.perf_counter() - t1) / 100
        del e
    loop_sec_per_step = sum(per_shape[tuple(s["shape"])] for s in specs)
    agg = n * steps / sec
    return dict(form=("persistent" if groups == 1 else "persistent, %d groups" % groups)
                if persistent else "stream branches",
                launches_per_step=len(hb.multi) if persistent else None,
                instances=n, size_classes=len(hb.shapes), steps=steps, sec=round(sec, 3),
                inst_steps_per_s=round(agg, 1), ms_per_step=round(sec / steps * 1e3, 3),
                plain_loop_inst_steps_per_s=round(n / loop_sec_per_step, 1),
                ratio_vs_plain_loop=round(agg / (n / loop_sec_per_step), 2),
                finite=finite, setup_sec=round(setup, 1))


def c2_input(nf=512, N=256, seed=1):
    """The microbench's coefficient fields: c[b, k, l] = g exp(-0.05 (k + l)) with
    g ~ N(0, 1) drawn by numpy's default_rng(seed), shape (nf, N, N)."""
    rng = np.random.default_rng(seed)
    k = np.arange(N, dtype=np.float64)
    decay = np.exp(-0.05 * (k[:, None] + k[None, :]))
    c = rng.standard_normal((nf, N, N))
    c *= decay
    return c


class C2Chain:
    """The microbench's kernel chain over a batch of N x N coefficient fields
    (device tensor c [nf, N, N]): K1-K2 backward Chebyshev transforms along x
    and y, K3-K4 forward transforms, K5-K6 ChebDirichlet from_chebyshev, K7-K8
    to_chebyshev, K9-K10 the Helmholtz right-hand side B (lambda = 0.01) along x
    and y, K11-K12 the four-diagonal solves along x then y (in place).
    stages: (name, fn, algorithmic bytes per field, MFMA flops per field or
    None); fn() runs one kernel on the buffers of the previous stage."""

    def __init__(self, c):
        import torch
        import pypde_amd
        from pypde_amd import ops
        from pypde_amd.templates import hholtz_matrices
        from pypde_amd.solver import Plan_fdma, PlanRHS
        nf, N, _ = c.shape
        M = N - 2
        self.c, self.nf, self.N = c, nf, N
        ch = pypde_amd.Base(N, "CH")
        cd = pypde_amd.Base(N, "CD")
        # the transforms in the kernel form tuned per shape: the parity-folded
        # MFMA product or the LDS-staged cosine sum (ops.ChebTransform)
        opB = ops.ChebTransform("bwd", N, ch.backward_matrix())
        opF = ops.ChebTransform("fwd", N, ch.forward_matrix())
        assert opB.mirror.mode and opF.mirror.mode
        s = cd.stencil_vector()
        Bm, _, Ax = hholtz_matrices(cd, 0.01, 1.0)
        rhs_x, rhs_y = PlanRHS(Bm, 2, 0), PlanRHS(Bm, 2, 1)
        lhs_x, lhs_y = Plan_fdma(Ax, 2, 0), Plan_fdma(Ax, 2, 1)
        b = self.bufs = {}
        b["u1"] = opB.apply(c, 0)
        b["u2"] = opB.apply(b["u1"], 1)
        b["c1"] = opF.apply(b["u2"], 0)
        b["c2"] = opF.apply(b["c1"], 1)
        fb = 8 * 2 * N * N
        fl = N * N * N  # MFMA work of a folded transform: half the dense 2 N^3
        self.stages = [
            ("K1_backward_x", lambda: opB.apply(c, 0, out=b["u1"]), fb, fl),
            ("K2_backward_y", lambda: opB.apply(b["u1"], 1, out=b["u2"]), fb, fl),
            ("K3_forward_x", lambda: opF.apply(b["u2"], 0, out=b["c1"]), fb, fl),
            ("K4_forward_y", lambda: opF.apply(b["c1"], 1, out=b["c2"]), fb, fl),
            ("K5_from_cheb_x", lambda: b.__setitem__("g1", cd.from_chebyshev(b["c2"], 0)),
             8 * (N * N + M * N), None),
            ("K6_from_cheb_y", lambda: b.__setitem__("g2", cd.from_chebyshev(b["g1"], 1)),
             8 * (M * N + M * M), None),
            ("K7_to_cheb_x", lambda: b.__setitem__("t1", ops.stencil_apply(s, b["g2"], 0)),
             8 * (M * M + N * M), None),
            ("K8_to_cheb_y", lambda: b.__setitem__("t2", ops.stencil_apply(s, b["t1"], 1)),
             8 * (N * M + N * N), None),
            ("K9_rhs_Bx", lambda: b.__setitem__("r1", rhs_x.solve(b["c2"])), 8 * (N * N + M * N),
             None),
            ("K10_rhs_By", lambda: b.__setitem__("r2", rhs_y.solve(b["r1"])), 8 * (M * N + M * M),
             None),
            ("K11_fdma_x", lambda: lhs_x.solve(b["x"]), 8 * 2 * M * M, None),
            ("K12_fdma_y", lambda: lhs_y.solve(b["x"]), 8 * 2 * M * M, None),
        ]
        self.bytes_per_field = sum(st[2] for st in self.stages)
        self.transforms = (opB, opF)

    def forms(self):
        """The kernel form of each transform stage (tuned per shape)."""
        opB, opF = self.transforms
        b = self.bufs
        return {"K1_backward_x": opB.form(self.c, 0), "K2_backward_y": opB.form(b["u1"], 1),
                "K3_forward_x": opF.form(b["u2"], 0), "K4_forward_y": opF.form(b["c1"], 1)}

    def prepare(self, name):
        """Set-up that is not part of a kernel: the in-place solves start from
        the right-hand side of K10."""
        if name == "K11_fdma_x":
            self.bufs["x"] = self.bufs["r2"].clone()

    def run_once(self):
        """Every kernel once, in chain order (the parity run)."""
        for name, fn, _, _ in self.stages:
            self.prepare(name)
            fn()


def microbench(nf=512, reps=10):
    """Kernel chain of 512 fields at 256x256 (C2Chain; per-field algorithmic bytes
    in its stage list); returns per-kernel time and HBM fraction.  Each launch is
    timed with HIP events enqueued behind a millisecond of GPU work, so that the
    events see the kernel and not the host's time to enqueue it."""
    import torch
    from pypde_amd import autotune, ops
    N = 256
    c = torch.as_tensor(c2_input(nf, N, seed=1)).to(ops.device())
    chain = C2Chain(c)
    out = {}

    def run(name, fn, bytes_per_field, flops=None):
        fn()  # warm
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            autotune.queue_delay()  # the GPU is busy while the host enqueues: events time the kernel
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) * 1e-3)
        t = float(np.median(ts))
        gbs = bytes_per_field * nf / t / 1e9
        rec = dict(us=round(t * 1e6, 2), GBps=round(gbs, 1), hbm_frac=round(gbs / HBM_PEAK_GBS, 4))
        if flops:
            rec["TFLOPps"] = round(flops * nf / t / 1e12, 2)
            rec["mfma_frac"] = round(flops * nf / t / 1e12 / FP64_PEAK_TFLOPS, 4)
        out[name] = rec

    # the measured roof: a device copy of the same nf x 256 x 256 array
    cp = torch.empty_like(c)
    run("device_copy", lambda: torch.mul(c, 1.0, out=cp), 8 * 2 * N * N)
    copy = out.pop("device_copy")
    del cp
    for name, fn, nbytes, flops in chain.stages:
        chain.prepare(name)
        run(name, fn, nbytes, flops)
    total_us = sum(v["us"] for v in out.values())
    total_bytes = chain.bytes_per_field * nf
    for v in out.values():
        v["copy_frac"] = round(v["GBps"] / copy["GBps"], 4)
    chain_gbs = total_bytes / (total_us * 1e-6) / 1e9
    forms = chain.forms()
    for k, f in forms.items():
        out[k]["form"] = "mfma" if f == "mfma" else "lds_dct/%d" % f
    return dict(fields=nf, N=N, kernels=out, chain_us=round(total_us, 1),
                chain_GBps=round(chain_gbs, 1),
                chain_hbm_frac=round(chain_gbs / HBM_PEAK_GBS, 4),
                chain_copy_frac=round(chain_gbs / copy["GBps"], 4),
                device_copy=dict(copy, kernel="torch.mul(x, 1.0, out=y), 512 x 256 x 256 float64"))


def transform_forms(sizes=(64, 128, 256, 512), reps=10):
    """The square Chebyshev transforms in both kernel forms (the parity-folded
    MFMA product and the LDS-staged cosine sums, ops.ChebTransform) at each N,
    over batches of about 32 M elements (512 fields of 256 x 256), both axes,
    backward and forward: microseconds per launch (median of reps, HIP events)
    and the form the tuner keeps."""
    import torch
    import pypde_amd
    from pypde_amd import autotune, ops
    out = {}
    for N in sizes:
        nf = max(1, (512 * 256 * 256) // (N * N))
        x = torch.as_tensor(c2_input(nf, N, seed=2)).to(ops.device())
        ch = pypde_amd.Base(N, "CH")
        res = {}
        for kind in ("bwd", "fwd"):
            op = ch._transform(kind)
            for axis in (0, 1):
                y = torch.empty_like(x)
                t = {}
                for form in op.FORMS:
                    op.apply(x, axis, out=y, form=form)
                    torch.cuda.synchronize()
                    ts = []
                    for _ in range(reps):
                        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                        autotune.queue_delay()
                        e0.record()
                        op.apply(x, axis, out=y, form=form)
                        e1.record()
                        torch.cuda.synchronize()
                        ts.append(e0.elapsed_time(e1) * 1e3)
                    t["mfma" if form == "mfma" else "lds_dct/%d" % form] = round(float(np.median(ts)), 2)
                chosen = op.form(x, axis)
                res["%s_axis%d" % (kind, axis)] = dict(
                    us=t, chosen="mfma" if chosen == "mfma" else "lds_dct/%d" % chosen,
                    hbm_frac_best=round(16.0 * x.numel() / (min(t.values()) * 1e-6) / 1e9 / HBM_PEAK_GBS, 4))
        out[str(N)] = dict(fields=nf, forms=res)
        del x
    return out


def time_variant(steps=100, N=256, **kw):
    """Steps/s of the fused N x N step (Ra=1e8, dt=1e-3) for another integrator /
    beta or size (graph replay)."""
    import torch
    from pypde_amd.navier import NavierStokes
    ns = NavierStokes(case="rbc", shape=(N, N), ra=1e8, pr=1.0, dt=1e-3, tsave=None, **kw)
    assert ns.fused
    ns.set_state(T=ns.T.forward(seeded_noise((N, N), seed=0)))
    ns.prepare_graphs()
    ns.update_many(5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ns.update_many(steps)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    return dict(kw, N=N, steps=steps, steps_per_s=round(steps / sec, 1),
                finite=bool(np.isfinite(ns.eval_Nu()[0])))


def time_steady(shape, ra, dt=0.01, tau=0.5, pre=300, tol=1e-10):
    """Time to a steady state by GPU Newton-Krylov: set_temperature(0.2), pre
    steps, then solve_steady_state(dt=tau) to max|F| <= tol."""
    import torch
    from pypde_amd.navier import NavierStokes
    ns = NavierStokes(case="rbc", shape=shape, ra=ra, pr=1.0, dt=dt, tsave=None)
    assert ns.fused
    ns.set_temperature(amplitude=0.2)
    ns.update_many(pre)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sol = ns.solve_steady_state(dt=tau, tol=tol, maxiter=60, disp=False)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    return dict(shape=list(shape), ra=ra, dt=dt, tau=tau, sec=round(sec, 3), success=bool(sol.success),
                nit=int(sol.nit), nfev=int(sol.nfev), fnorm=float(sol.fnorm),
                nu_plate=float(ns.eval_Nu()[0]))


def time_adjoint_descent(N, steps=20, ra=1e4, dt=0.005):
    """Adjoint-descent updates per second (navier.rbc2d_adj drop-in) at N x N,
    replayed from the update's HIP graph and, for comparison, op by op."""
    import torch
    from pypde_amd.navier import NavierStokesAdjoint
    out = dict(N=N, ra=ra, dt=dt, steps=steps)
    for use_graph in (True, False):
        nsa = NavierStokesAdjoint(case="rbc", shape=(N, N), ra=ra, pr=1.0, dt=dt, tsave=None,
                                  use_graph=use_graph)
        nsa.set_temperature(amplitude=0.2)
        for _ in range(3):
            nsa.update()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            nsa.update()
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        key = "ms_per_update" if use_graph else "ms_per_update_op_by_op"
        out[key] = round(sec / steps * 1e3, 3)
        if use_graph:
            out["finite"] = bool(torch.isfinite(nsa.T.vhat).all())
    return out


def time_adjoint(N, nsteps, ckpt=100, ra=5e4, dt=5e-3):
    """Whole gradient dJ/dT0 of J = mean_k Nu_plate(T_k) over nsteps: forward run
    with checkpoints, segment recomputation and the reverse VJP sweep."""
    import torch
    from pypde_amd.adjoint import Trajectory
    from pypde_amd.engine import RBCBatch
    from pypde_amd.field import Field
    from pypde_amd.bases import Base
    eng = RBCBatch((N, N), [dict(ra=ra, pr=1.0, dt=dt)], case="rbc")
    Tf = Field([Base(N, "CN"), Base(N, "CD")])
    X0 = torch.zeros_like(eng.X)
    X0[0, 2].copy_(Tf.forward(seeded_noise((N, N), seed=0)))
    pres0 = torch.zeros_like(eng.pres)
    traj = Trajectory(eng, ckpt=ckpt)
    traj.grad_nusselt(X0, pres0, min(nsteps, 2 * ckpt))  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J, Xb, pb = traj.grad_nusselt(X0, pres0, nsteps)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    return dict(N=N, steps=nsteps, checkpoint_every=ckpt, ra=ra, dt=dt, sec=round(sec, 3),
                steps_per_s=round(nsteps / sec, 2), J=float(J[0]),
                grad_norm=float(torch.linalg.norm(Xb[0, 2])),
                finite=bool(torch.isfinite(Xb).all().item()))


def cpu_baseline(steps=6, warm=1):
    """Oracle (CPU restatement of the reference, scipy DCT-I + C ports of the
    Fortran kernels, per-call dense Poisson formation as the reference) at
    256x256, one thread."""
    from threadpoolctl import threadpool_limits
    from oracle import pypde_oracle as o
    N = 256
    with threadpool_limits(limits=1):
        ns = o.NavierStokes(case="rbc", shape=(N, N), ra=1e8, pr=1.0, dt=1e-3, tsave=None)
        o.seeded_initial_state(ns, seed=0)
        for _ in range(warm):
            ns.update()
        t0 = time.perf_counter()
        for _ in range(steps):
            ns.update()
        sec = time.perf_counter() - t0
    return dict(value=round(steps / sec, 3), unit="steps/s", cores=1, kind="port",
                sample="%d Euler steps of the 256x256 rbc workload (Ra=1e8, dt=1e-3) after %d warm-up" %
                       (steps, warm), sec=round(sec, 2), host=host_cpu())


def gemm_k_sweep(m=384, n=192, batch=18, ks=(16, 32, 64, 127, 254, 508), variants=(2, 1, 27, 33, 40)):
    """Time of a plain strided-batch GEMM at the y-pass's shape (18 products of
    384 x 192 over K, B transposed) against K, for a few kernel variants: the
    intercept is the launch's fixed cost (dispatch, ramp, epilogue, tail), the
    slope the cost per K slab.  Each launch timed by the tuner's timing (five
    back-to-back launches behind a queued millisecond, best of three)."""
    import ctypes
    import torch
    from pypde_amd import autotune
    from pypde_amd._lib import BIdx, GemmDesc, check, lib

    def bidx(s1):
        b = BIdx()
        b.s1, b.d1, b.m1, b.s2, b.d2, b.m2 = s1, 1, 0, 0, 1, 0
        return b
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = {}
    for k in ks:
        A = torch.randn(batch, m, k, dtype=torch.float64, device="cuda")
        Bm = torch.randn(batch, n, k, dtype=torch.float64, device="cuda")
        C = torch.empty(batch, m, n, dtype=torch.float64, device="cuda")
        d = GemmDesc()
        d.transA, d.transB, d.m, d.n, d.k, d.batch = 0, 1, m, n, k, batch
        d.alpha, d.beta = 1.0, 0.0
        d.A, d.B, d.C = A.data_ptr(), Bm.data_ptr(), C.data_ptr()
        d.lda, d.ldb, d.ldc = k, k, n
        d.iA, d.iB, d.iC = bidx(m * k), bidx(n * k), bidx(m * n)
        d.alpha_div = d.gamma_div = 1
        d.ska = d.skb = d.csn = 1
        rec = {}
        for v in variants:
            d.variant = v
            t = autotune.time_candidates(
                lambda _: check(lib.pypde_dgemm_batched(ctypes.byref(d), st), "gemm"), [0])
            rec[str(v)] = round(t[0], 2)
        ref = torch.bmm(A, Bm.transpose(1, 2).contiguous())
        err = float((C - ref).abs().max() / ref.abs().max())
        out[str(k)] = dict(us=rec, rel_err=err)
        print("gemm K sweep: K=%d done" % k, file=sys.stderr, flush=True)
    return out


def gemm_variants():
    """Every plain GEMM kernel variant on each plain GEMM launch of the 256x256
    step (single descriptors and grouped launches; the folded x-pass is not a
    plain GEMM), timed like the tuner does; for the y-pass also with its A
    operand read contiguously (ska = 1: the same shape without the
    every-other-column access).  The engine state is restored afterwards."""
    import ctypes
    import torch
    from pypde_amd import autotune
    from pypde_amd._lib import GemmDesc, check, lib
    from pypde_amd.engine import _Gemm, _GemmGroup, _SplitRhs
    from pypde_amd.navier import NavierStokes
    ns = NavierStokes(case="rbc", shape=(256, 256), ra=1e8, pr=1.0, dt=1e-3, tsave=None)
    eng = ns.engine
    eng.step_eager()
    torch.cuda.synchronize()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    saved = [t.clone() for t in (eng.X, eng.pres, eng.P)]

    def sweep(descs):
        arr = (GemmDesc * len(descs))(*descs)

        def run(v):
            for i in range(len(descs)):
                arr[i].variant = v
            if len(descs) == 1:
                check(lib.pypde_dgemm_batched(ctypes.byref(arr), s), "variant sweep")
            else:
                check(lib.pypde_dgemm_grouped(ctypes.byref(arr), len(descs), s), "variant sweep")
        return autotune.time_candidates(run, autotune.GEMM_CANDIDATES)
    out = {}
    for l in eng.launches:
        for p in ([l.group] if isinstance(l, _SplitRhs) else [l]):
            if isinstance(p, _GemmGroup):
                descs = [GemmDesc.from_buffer_copy(m.desc) for m in p.members]
            elif isinstance(p, _Gemm) and p.desc.fold == 0:
                descs = [GemmDesc.from_buffer_copy(p.desc)]
            else:
                continue
            t = sweep(descs)
            ranked = sorted(t.items(), key=lambda kv: kv[1])
            e = dict(in_step=descs[0].variant, us_in_step=round(t[descs[0].variant], 2),
                     ranked=[[v, round(u, 2)] for v, u in ranked[:12]])
            if any(d.ska == 2 for d in descs):
                for d in descs:
                    d.ska = 1
                t1 = sweep(descs)
                b1 = min(t1, key=t1.get)
                e["contiguous_A"] = dict(fastest=b1, us=round(t1[b1], 2),
                                         us_in_step_variant=round(t1[e["in_step"]], 2))
            out[p.name] = e
            print("variants: %s done" % p.name, file=sys.stderr, flush=True)
    for t, v in zip((eng.X, eng.pres, eng.P), saved):
        t.copy_(v)
    torch.cuda.synchronize()
    return out


def batched_launches(nb=64):
    """Per-launch times of the batched 256x256 step (nb instances, the batch
    of time_batched): each launch replayed 20 times back to back from a HIP
    graph, with its bound and its achieved rate."""
    import torch
    from pypde_amd.engine import RBCBatch, _GemmGroup, _SplitRhs
    N = 256
    rng = np.random.default_rng(2024)
    ras = 10.0 ** rng.uniform(4, 7, size=nb)
    dts = np.clip(0.01 * np.sqrt(1e5 / ras), 1e-3, 1e-2)
    eng = RBCBatch((N, N), [dict(ra=float(r), pr=1.0, dt=float(d)) for r, d in zip(ras, dts)],
                   case="rbc")
    eng.capture(1)
    eng.step(2)
    torch.cuda.synchronize()
    out, total = [], 0.0
    for l in eng.launches:
        for part in ([l.group, l.reduce] if isinstance(l, _SplitRhs) else [l]):
            t = launch_graph_time(eng, part)
            bound, amount = launch_cost(part.name, part, eng)
            rate = amount / t / (1e12 if bound == "mfma" else 1e9)
            out.append(dict(launch=part.name, us=round(t * 1e6, 2), bound=bound,
                            rate=round(rate, 2), unit="TFLOP/s" if bound == "mfma" else "GB/s"))
            total += t
    return dict(instances=nb, launches=out, sum_us=round(total * 1e6, 1))


def rhs_forms():
    """The banded right-hand side of the step in each kernel form (one output
    per thread, row strips of 4 and 8, LDS tiles), microseconds per launch
    (20 back-to-back launches, HIP events), at the workloads' shapes."""
    import ctypes
    import torch
    from pypde_amd._lib import RbcRhs, lib, check
    from pypde_amd.engine import RBCBatch
    from pypde_amd import autotune
    out = {}
    for shape, B in (((256, 256), 1), ((256, 256), 64), ((128, 128), 16), ((64, 64), 16),
                     ((192, 192), 16), ((512, 512), 1)):
        params = [dict(ra=1e6, pr=1.0, dt=1e-3)] * B
        eng = RBCBatch(shape, params, case="rbc", group_launches=False)
        a = None
        for l in eng.launches:
            for arg in getattr(l, "args", ()):
                if isinstance(getattr(arg, "_obj", None), RbcRhs):
                    a = RbcRhs.from_buffer_copy(arg._obj)
        st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        res = {}
        for rs in (1, 4, 8, 16):
            a.rows_per_thread = rs
            check(lib.pypde_rbc_rhs_banded(ctypes.byref(a), st), "rhs form")
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            autotune.queue_delay()
            e0.record()
            for _ in range(20):
                lib.pypde_rbc_rhs_banded(ctypes.byref(a), st)
            e1.record()
            torch.cuda.synchronize()
            res[rs] = round(e0.elapsed_time(e1) * 1e3 / 20, 2)
        out["%dx%d_B%d" % (shape[0], shape[1], B)] = res
        del eng
    return out


def hetero_launches():
    """Per-launch times of the mixed-size batch (bench.c4_specs): each multi
    launch of the persistent form, and each size class's launches in the
    stream-branch form, timed alone (20 back-to-back launches, HIP events)."""
    import ctypes
    import torch
    from pypde_amd.batch import HeteroBatch
    from pypde_amd import autotune

    def ev(fn, reps=20):
        st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        fn(st)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        autotune.queue_delay()
        e0.record()
        for _ in range(reps):
            fn(st)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) * 1e3 / reps
    specs = c4_specs()
    seeds = [1000 + i for i in range(len(specs))]
    hb = HeteroBatch(specs, case="rbc", persistent=True)
    hb.seed(seeds, amplitude=1e-3)
    saved = [[t.clone() for t in (e.X, e.pres, e.P)] for e in hb.classes]
    pers = [[m.name, m.kind, round(ev(m), 2)] for m in hb.multi]
    for e, sv in zip(hb.classes, saved):
        for t, v in zip((e.X, e.pres, e.P), sv):
            t.copy_(v)
    out = dict(persistent=pers, persistent_sum_us=round(sum(r[2] for r in pers), 1))
    del hb
    hb = HeteroBatch(specs, case="rbc", persistent=False)
    hb.seed(seeds, amplitude=1e-3)
    per = {}
    for e in hb.classes:
        saved = [t.clone() for t in (e.X, e.pres, e.P)]
        for l in e.launches:
            per.setdefault(l.name, []).append(round(ev(l), 2))
        for t, v in zip((e.X, e.pres, e.P), saved):
            t.copy_(v)
    out["branches"] = {k: dict(sum_us=round(sum(v), 1), per_class_us=v) for k, v in per.items()}
    out["branches_sum_us"] = round(sum(sum(v) for v in per.values()), 1)
    out["classes"] = [dict(shape=list(sh), instances=e.B) for sh, e in zip(hb.shapes, hb.classes)]
    return out


def cpu_table(threads=(1, None)):
    """The oracle on the host at every workload (one thread, then every core
    through threadpoolctl; scipy's FFTs stay single-threaded as the
    reference's): 64x64 (Ra=1e5, dt=0.01) over its full 1000 steps; the
    512 x 256x256 transform/solve chain on a sample of fields, scaled to 512;
    512x512 (Ra=1e8, dt=1e-3) 10 timed steps after 2; the mixed-size batch on
    one instance per size class, scaled to the 256 instances; 128x128
    (Ra=1e6, dt=1e-3) 20 steps after 2; 256x256 (Ra=1e8, dt=1e-3) 20 steps
    after 5."""
    from threadpoolctl import threadpool_limits
    from oracle import pypde_oracle as o

    def steps_rate(shape, ra, dt, steps, warm, seed=0):
        ns = o.NavierStokes(case="rbc", shape=shape, ra=ra, pr=1.0, dt=dt, tsave=None)
        o.seeded_initial_state(ns, seed=seed)
        for _ in range(warm):
            ns.update()
        t0 = time.perf_counter()
        for _ in range(steps):
            ns.update()
        return steps / (time.perf_counter() - t0)

    def chain_rate(nsample=16, nf=512, N=256):
        c = c2_input(nsample, N, seed=1)
        ch, cd = o.Basis(N, "CH"), o.Basis(N, "CD")
        Bm, _, Ax = o.hholtz_matrices(cd, 0.01, 1.0)
        along = o._along
        mx, my, fx, fy = o.MatAxis(Bm, 0), o.MatAxis(Bm, 1), o.FdmaAxis(Ax, 0), o.FdmaAxis(Ax, 1)
        t0 = time.perf_counter()
        for f in range(nsample):
            u1 = along(ch._cheb_backward, c[f], 0)
            u2 = along(ch._cheb_backward, u1, 1)
            c1 = along(ch._cheb_forward, u2, 0)
            c2 = along(ch._cheb_forward, c1, 1)
            g1 = along(cd.from_chebyshev, c2, 0)
            g2 = along(cd.from_chebyshev, g1, 1)
            along(cd.to_chebyshev, along(cd.to_chebyshev, g2, 0), 1)
            fy(fx(my(mx(c2))))
        return nsample / (time.perf_counter() - t0)

    out = {}
    for th in threads:
        label = "1_thread" if th == 1 else "all_cores"
        r = {}
        with threadpool_limits(limits=th):
            t0 = time.perf_counter()
            r["c1_64x64_1000_steps"] = dict(steps_per_s=round(steps_rate((64, 64), 1e5, 0.01, 1000, 0), 2))
            print("cpu table %s: c1 done (%.0f s)" % (label, time.perf_counter() - t0), file=sys.stderr, flush=True)
            r["c2_chain_fields_per_s"] = dict(
                value=round(chain_rate(), 2), sample="16 of the 512 fields, scaled",
                batch_s=None)
            r["c2_chain_fields_per_s"]["batch_s"] = round(512 / r["c2_chain_fields_per_s"]["value"], 2)
            print("cpu table %s: c2 done" % label, file=sys.stderr, flush=True)
            r["c3_512x512"] = dict(steps_per_s=round(steps_rate((512, 512), 1e8, 1e-3, 10, 2), 3),
                                   sample="10 timed steps after 2")
            print("cpu table %s: c3 done" % label, file=sys.stderr, flush=True)
            specs = c4_specs()
            classes = {}
            for i, sp in enumerate(specs):
                classes.setdefault(sp["shape"], []).append(i)
            sec_per_step = {}
            for shape, idx in classes.items():
                sp = specs[idx[0]]
                sec_per_step[shape] = 1.0 / steps_rate(shape, sp["ra"], sp["dt"], 4, 1, seed=1000 + idx[0])
            total = sum(sec_per_step[sp["shape"]] for sp in specs)
            r["c4_mixed_256"] = dict(instance_steps_per_s=round(len(specs) / total, 2),
                                     sample="one instance per size class (%d classes), 4 steps "
                                            "after 1, scaled to the 256 instances" % len(classes))
            print("cpu table %s: c4 done" % label, file=sys.stderr, flush=True)
            r["c5_128x128_forward"] = dict(steps_per_s=round(steps_rate((128, 128), 1e6, 1e-3, 20, 2), 2))
            r["metric_256x256"] = dict(steps_per_s=round(steps_rate((256, 256), 1e8, 1e-3, 20, 5), 3))
            r["sec"] = round(time.perf_counter() - t0, 1)
            print("cpu table %s: done (%.0f s)" % (label, r["sec"]), file=sys.stderr, flush=True)
        out[label] = r
    out["host"] = host_cpu()
    return out


def host_cpu():
    """Model name and logical CPU count of the host the CPU baseline ran on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(model=model, logical_cpus=os.cpu_count())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip batched and microbench")
    ap.add_argument("--cpu-steps", type=int, default=150)
    ap.add_argument("--micro", action="store_true", help="run the microbench only")
    ap.add_argument("--gemm-variants", action="store_true",
                    help="time every plain GEMM variant on the 256x256 step's GEMM launches")
    ap.add_argument("--gemm-k-sweep", action="store_true",
                    help="time a y-pass-shaped GEMM against K (fixed cost and cost per K slab)")
    ap.add_argument("--batched-launches", action="store_true",
                    help="per-launch times of the 64 x 256x256 batched step")
    ap.add_argument("--rhs-forms", action="store_true",
                    help="time the banded right-hand side's kernel forms at the workloads' shapes")
    ap.add_argument("--hetero-launches", action="store_true",
                    help="per-launch times of the mixed-size batch in both forms")
    ap.add_argument("--cpu-table", action="store_true",
                    help="time the oracle on the host at every workload (one thread and all cores)")
    ap.add_argument("--transforms", action="store_true",
                    help="time both kernel forms of the square Chebyshev transforms at "
                         "N = 64, 128, 256, 512")
    ap.add_argument("--adjoint", action="store_true",
                    help="run the adjoint gradient at 256x256 (200 steps) only")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--all-launches", action="store_true",
                    help="after the plain step, replay every launch of it 21 times in step "
                         "order (one block per launch, for per-launch PMC counts)")
    ap.add_argument("--tune-db", metavar="PATH",
                    help="tune every workload of this benchmark from scratch and write the "
                         "kernel decisions to PATH (the database pypde_amd/tune_db.json)")
    ap.add_argument("--tune-db-update", metavar="PATH",
                    help="as --tune-db, but keep the stored decisions and tune only the "
                         "shapes they do not cover")
    args = ap.parse_args()
    if args.tune_db:
        os.environ["PYPDE_TUNE_DB"] = ""  # measure, do not load stored decisions
    elif args.tune_db_update:
        args.tune_db = args.tune_db_update
    if args.gpus != 1:
        print("bench.py: pypde_amd runs one instance set on one GPU (got --gpus %d); "
              "multi-GPU execution is not implemented." % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.micro:
        print(json.dumps({"microbench": microbench()}))
        return
    if args.transforms:
        print(json.dumps({"transforms": transform_forms()}))
        return
    if args.cpu_table:
        print(json.dumps({"cpu_table": cpu_table()}))
        return
    if args.gemm_variants:
        print(json.dumps({"gemm_variants": gemm_variants()}))
        return
    if args.gemm_k_sweep:
        print(json.dumps({"gemm_k_sweep": gemm_k_sweep()}))
        return
    if args.batched_launches:
        print(json.dumps({"batched_launches": batched_launches()}))
        return
    if args.rhs_forms:
        print(json.dumps({"rhs_forms": rhs_forms()}))
        return
    if args.hetero_launches:
        print(json.dumps({"hetero_launches": hetero_launches()}))
        return
    if args.adjoint:
        print(json.dumps({"adjoint": time_adjoint(256, 200, ckpt=100)}))
        return
    if args.tune_db:
        from pypde_amd import autotune
        args.steps, args.warmup = 20, 2

        def stage(name, fn, *a, **kw):
            t0 = time.perf_counter()
            fn(*a, **kw)
            print("tuned %s (%.1f s, %d decisions)" % (name, time.perf_counter() - t0,
                                                       len(autotune.cache())),
                  file=sys.stderr, flush=True)
        stage("plain", time_plain, args)
        stage("batched", time_batched, args)
        stage("hetero", time_hetero)
        stage("hetero_groups", time_hetero, groups=2)
        stage("hetero_groups4", time_hetero, groups=4)
        stage("hetero_branches", time_hetero, persistent=False)
        stage("rk3", time_variant, steps=5, integrator="rk3")
        stage("beta", time_variant, steps=5, beta=0.5)
        stage("512", time_variant, steps=5, N=512)
        stage("microbench", microbench)
        stage("transforms", transform_forms)
        for shape, ra in (((20, 16), 1e4), ((64, 64), 1e4), ((128, 128), 5e4)):
            stage("steady %s" % (shape,), time_steady, shape, ra, pre=20, tol=1e-3)
        stage("adjoint descent", time_adjoint_descent, 256, steps=2)
        stage("adjoint 256", time_adjoint, 256, 20, ckpt=10)
        stage("adjoint 128", time_adjoint, 128, 20, ckpt=10)
        n = autotune.save_db(args.tune_db)
        print(json.dumps({"tune_db": args.tune_db, "entries": n}))
        return
    plain = time_plain(args)
    if args.all_launches:
        from pypde_amd.engine import _GemmGroup, _SplitRhs
        eng = plain["engine"]
        order = []
        for launch in eng.launches:
            parts = [launch.group, launch.reduce] if isinstance(launch, _SplitRhs) else [launch]
            for l in parts:
                launch_graph_time(eng, l)
                kernels = len(l.members) if isinstance(l, _GemmGroup) and l.split else 1
                order.append([l.name, kernels])
        print(json.dumps({"launch_order": order}))
    t_start = time.perf_counter()

    def done(what):
        print("bench: %s done (%.0f s)" % (what, time.perf_counter() - t_start), file=sys.stderr,
              flush=True)
    done("plain")
    extra = {}
    if not args.no_extra:
        b = time_batched(args)
        extra["batched"] = dict(b, ratio_vs_plain=round(b["inst_steps_per_s"] / plain["steps_per_s"], 2))
        done("batched")
        extra["hetero"] = time_hetero()
        extra["hetero_groups"] = [time_hetero(groups=g) for g in (2, 4)]
        extra["hetero_branches"] = time_hetero(persistent=False)
        done("hetero")
        extra["variants"] = [time_variant(integrator="rk3"), time_variant(beta=0.5),
                             time_variant(N=512)]
        done("variants")
        extra["microbench"] = microbench()
        extra["transforms"] = transform_forms()
        done("microbench")
        extra["steady"] = [time_steady((20, 16), 1e4), time_steady((64, 64), 1e4),
                           time_steady((128, 128), 5e4)]
        done("steady")
        extra["adjoint_descent"] = time_adjoint_descent(256)
        # checkpoints every 100 steps (the configured run), and every state kept
        # in HBM (400 MB at 256x256 over 200 steps), which removes the recomputation
        extra["adjoint"] = {"256x256": time_adjoint(256, 200), "128x128": time_adjoint(128, 2000),
                            "256x256_all_states": time_adjoint(256, 200, ckpt=1)}
        done("adjoint")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(steps=args.cpu_steps)
        done("cpu baseline")
    line = {
        "metric": "RBC2D FP64 time-steps/sec @256x256 (plain vs batched vs adjoint), % HBM roofline",
        "value": round(plain["steps_per_s"], 2),
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(plain["sec"] / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "rbc2d_plain_256x256", "shape": [256, 256], "case": "rbc",
                   "ra": 1e8, "pr": 1.0, "dt": 1e-3, "integrator": "euler", "dealias": 1.5,
                   "init": "conductive + uniform noise 1e-3, seed 0", "graph_replay": True},
        "roofline": {"bound": plain["bound"], "achieved": round(plain["achieved"], 3),
                     "peak": plain["peak"], "unit": plain["unit"],
                     "frac": round(plain["achieved"] / plain["peak"], 5),
                     "traffic": TRAFFIC_BYTES.get(plain["dominant"]),
                     "kernel": plain["dominant"], "avg_us": round(plain["avg_s"] * 1e6, 3)},
        "cpu_baseline": cpu,
        "kernels_us": plain["kernels"],
        "launch_variants": plain["engine"].variants(),
        "eager_ms_per_step": round(plain["eager_sec"] / args.steps * 1e3, 4),
        "nu_final": plain["nu"],
        "finite": plain["finite"],
    }
    line["adjoint"] = None
    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
