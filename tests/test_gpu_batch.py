"""Heterogeneous batches: instances of different shapes, Rayleigh numbers and
time steps advanced by one graph launch per step must equal the oracle's
single-instance runs (reference navier/rbc2d.py:396-434) of each instance."""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


def host(t):
    return t.detach().cpu().numpy()


def test_hetero_batch_matches_oracle_per_instance():
    from pypde_amd.batch import HeteroBatch, default_dt
    from oracle import pypde_oracle as o
    specs = [dict(shape=(20, 16), ra=1e4, pr=1.0), dict(shape=(24, 20), ra=3e5, pr=1.0),
             dict(shape=(16, 24), ra=2e6, pr=0.7), dict(shape=(20, 16), ra=5e5, pr=1.0),
             dict(shape=(24, 20), ra=1e5, pr=1.0)]
    for s in specs:
        s["dt"] = default_dt(s["ra"])
    hb = HeteroBatch(specs, case="rbc")
    assert len(hb.classes) == 3
    seeds = [100 + i for i in range(len(specs))]
    hb.seed(seeds, amplitude=0.05)
    refs = []
    for s, sd in zip(specs, seeds):
        r = o.NavierStokes(case="rbc", shape=s["shape"], ra=s["ra"], pr=s["pr"], dt=s["dt"],
                           tsave=None)
        o.seeded_initial_state(r, seed=sd, amplitude=0.05)
        refs.append(r)
    for i, r in enumerate(refs):
        assert rel(host(hb.state(i)["T"]), r.T.vhat) < 1e-13, i
    hb.capture(1)
    nsteps = 5
    hb.step(nsteps)
    for r in refs:
        for _ in range(nsteps):
            r.update()
    for i, r in enumerate(refs):
        st = hb.state(i)
        for f in ("U", "V", "T"):
            e = rel(host(st[f]), getattr(r, f).vhat)
            assert e < 1e-11, (i, f, e)
        assert rel(host(st["pres"]), r.pres.vhat) < 1e-11, i
    nu = hb.nusselt()
    for i, r in enumerate(refs):
        assert abs(nu[i] - r.eval_Nu()[0]) < 1e-10 * max(1.0, abs(nu[i])), i
    assert hb.finite()


def test_default_dt_rule():
    from pypde_amd.batch import default_dt
    assert default_dt(1e5) == pytest.approx(0.01)
    assert default_dt(1e7) == pytest.approx(1e-3)
    assert default_dt(1e8) == pytest.approx(1e-3)
    assert default_dt(1e4) == pytest.approx(1e-2)


def test_persistent_hetero_batch_matches_oracle_per_instance():
    """The persistent form (every operator of the step one multi launch for all
    size classes) against the oracle's single-instance runs, as above."""
    from pypde_amd.batch import HeteroBatch, default_dt
    from oracle import pypde_oracle as o
    specs = [dict(shape=(20, 16), ra=1e4, pr=1.0), dict(shape=(24, 20), ra=3e5, pr=1.0),
             dict(shape=(16, 24), ra=2e6, pr=0.7), dict(shape=(20, 16), ra=5e5, pr=1.0),
             dict(shape=(24, 20), ra=1e5, pr=1.0)]
    for s in specs:
        s["dt"] = default_dt(s["ra"])
    hb = HeteroBatch(specs, case="rbc", persistent=True)
    assert len(hb.classes) == 3 and len(hb.multi) == 13
    seeds = [100 + i for i in range(len(specs))]
    hb.seed(seeds, amplitude=0.05)
    hb.capture(1)
    nsteps = 5
    hb.step(nsteps)
    for i, (s, sd) in enumerate(zip(specs, seeds)):
        r = o.NavierStokes(case="rbc", shape=s["shape"], ra=s["ra"], pr=s["pr"], dt=s["dt"],
                           tsave=None)
        o.seeded_initial_state(r, seed=sd, amplitude=0.05)
        for _ in range(nsteps):
            r.update()
        st = hb.state(i)
        for f in ("U", "V", "T"):
            e = rel(host(st[f]), getattr(r, f).vhat)
            assert e < 1e-11, (i, f, e)
        assert rel(host(st["pres"]), r.pres.vhat) < 1e-11, i


def test_persistent_c4_batch_matches_branch_form():
    """The 256-instance mixed-size batch (N per axis from {64, 96, 128, 192}):
    13 launches per step for its 16 size classes, and after 5 steps every
    instance equal to the stream-branch form within 1e-10 (the two forms may
    run different GEMM tiles)."""
    import sys
    import os
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    from pypde_amd.batch import HeteroBatch
    specs = bench.c4_specs()
    seeds = [1000 + i for i in range(len(specs))]
    out = []
    for persistent in (True, False):
        hb = HeteroBatch(specs, case="rbc", persistent=persistent)
        if persistent:
            assert len(hb.classes) == 16 and len(hb.multi) == 13
        hb.seed(seeds, amplitude=1e-3)
        hb.capture(1)
        hb.step(5)
        out.append([{f: host(t) for f, t in hb.state(i).items()} for i in range(len(specs))])
        del hb
    worst = 0.0
    for a, b in zip(*out):
        for f in ("U", "V", "T", "pres"):
            worst = max(worst, rel(a[f], b[f]))
    print("persistent vs branch form, worst relative difference %.2e" % worst)
    assert worst < 1e-10


def test_persistent_batch_plan_through_capi_equals_its_launches():
    """pypde_rbc_step_batched (the persistent step's multi launches recorded as
    one step plan, run through ctypes) gives bit for bit what the same launches
    give one by one, over three steps."""
    import ctypes
    import torch
    from pypde_amd._lib import lib
    from pypde_amd.batch import HeteroBatch, default_dt
    specs = [dict(shape=(20, 16), ra=1e4, pr=1.0), dict(shape=(24, 20), ra=3e5, pr=1.0),
             dict(shape=(16, 24), ra=2e6, pr=0.7)]
    for s in specs:
        s["dt"] = default_dt(s["ra"])
    seeds = [7, 8, 9]
    hb = HeteroBatch(specs, case="rbc", persistent=True)
    hb.seed(seeds, amplitude=0.05)
    plan = hb.build_plan()
    assert lib.pypde_plan_size(plan.handle) == len(hb.multi) == 13
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib.pypde_rbc_step_batched(plan.handle, 3, s) == 0
    torch.cuda.synchronize()
    via_plan = [{f: host(t) for f, t in hb.state(i).items()} for i in range(len(specs))]
    hb.seed(seeds, amplitude=0.05)
    for _ in range(3):
        for m in hb.multi:
            m(s)
    torch.cuda.synchronize()
    for i in range(len(specs)):
        for f, a in via_plan[i].items():
            assert np.array_equal(a, host(hb.state(i)[f])), (i, f)
    assert lib.pypde_rbc_step_batched(None, 1, s) != 0
    assert lib.pypde_rbc_step_batched(plan.handle, -1, s) != 0


def test_persistent_groups_match_single_chain():
    """The persistent form with its size classes in two groups (two chains of
    multi launches on parallel graph branches) gives what the single chain
    gives, to rounding: each class runs the same operators, but the tuner may
    choose another GEMM variant for a launch over fewer members."""
    from pypde_amd.batch import HeteroBatch, default_dt
    specs = [dict(shape=(20, 16), ra=1e4, pr=1.0), dict(shape=(24, 20), ra=3e5, pr=1.0),
             dict(shape=(16, 24), ra=2e6, pr=0.7), dict(shape=(20, 16), ra=5e5, pr=1.0)]
    for s in specs:
        s["dt"] = default_dt(s["ra"])
    seeds = [50 + i for i in range(len(specs))]
    out = []
    for groups in (1, 2):
        hb = HeteroBatch(specs, case="rbc", persistent=True, groups=groups)
        assert len(hb.group_multi) == groups and len(hb.multi) == 13 * groups
        hb.seed(seeds, amplitude=0.05)
        hb.capture(1)
        hb.step(4)
        out.append([{f: host(t) for f, t in hb.state(i).items()} for i in range(len(specs))])
    for a, b in zip(*out):
        for f in a:
            assert np.allclose(a[f], b[f], rtol=0, atol=1e-11 * max(1.0, np.abs(a[f]).max())), f
