"""GPU parity of the RBC2D time step (fused engine and op-by-op path) against
the reference's own vectors (tests/golden) and the oracle."""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

CASES = {
    "linear_eu": dict(case="linear", integrator="eu", beta=1.0),
    "rbc_eu": dict(case="rbc", integrator="eu", beta=1.0),
    "zero_eu": dict(case="zero", integrator="eu", beta=1.0),
    "rbc_rk3": dict(case="rbc", integrator="rk3", beta=1.0),
    "rbc_eu_beta": dict(case="rbc", integrator="eu", beta=0.5),
}
# Tolerances (norm-wise relative): the fused step regroups the reference's
# operator chains into dense products, so sums are reordered; one step stays
# within 1e-12 of the reference.
TOL_STEP = 1e-12


def host(t):
    return t.detach().cpu().numpy()


def make(name, use_graph=True):
    from pypde_amd.navier import NavierStokes
    kw = dict(CASES[name])
    case = kw.pop("case")
    return NavierStokes(case=case, shape=(20, 16), ra=1e5, pr=1.0, dt=0.01, tsave=None,
                        dealias=True, aspect=1.0, use_graph=use_graph, **kw)


@pytest.mark.parametrize("name", sorted(CASES))
def test_steps_match_reference(golden, name):
    """Every variant (Euler / RK3, beta = 1 / 0.5, all cases) runs on the fused engine."""
    ns = make(name)
    assert ns.fused, name
    pre = "step_%s_" % name
    ns.set_state(T=golden[pre + "in_T"], U=golden[pre + "in_U"], V=golden[pre + "in_V"],
                 P=golden[pre + "in_P"], pres=golden[pre + "in_pres"])
    for n in range(3):
        ns.update()
        if n == 0:
            for f in ("T", "U", "V", "pres"):
                e = rel(host(getattr(ns, f).vhat), golden[pre + "out1_" + f])
                assert e < TOL_STEP, (f, e)
    for f in ("T", "U", "V", "pres"):
        e = rel(host(getattr(ns, f).vhat), golden[pre + "out3_" + f])
        assert e < 3 * TOL_STEP, (f, e)


def test_graph_replay_matches_eager(golden):
    a, b = make("rbc_eu", use_graph=True), make("rbc_eu", use_graph=False)
    pre = "step_rbc_eu_in_"
    for ns in (a, b):
        ns.set_state(T=golden[pre + "T"], U=golden[pre + "U"], V=golden[pre + "V"],
                     pres=golden[pre + "pres"])
    for _ in range(5):
        a.update()
        b.update()
    for f in ("T", "U", "V", "pres", "P"):
        assert np.array_equal(host(getattr(a, f).vhat), host(getattr(b, f).vhat)), f


def test_multi_step_graph_matches_single_steps(golden):
    """update_many replays a graph of several steps plus single steps; the
    state equals that of the same number of single updates bit for bit."""
    a, b = make("rbc_eu", use_graph=True), make("rbc_eu", use_graph=True)
    assert a.GRAPH_BLOCK > 1
    pre = "step_rbc_eu_in_"
    for ns in (a, b):
        ns.set_state(T=golden[pre + "T"], U=golden[pre + "U"], V=golden[pre + "V"],
                     pres=golden[pre + "pres"])
    n = 2 * a.GRAPH_BLOCK + 3
    a.update_many(n)
    assert a.GRAPH_BLOCK in a.engine.graphs
    for _ in range(n):
        b.update()
    for f in ("T", "U", "V", "pres", "P"):
        assert np.array_equal(host(getattr(a, f).vhat), host(getattr(b, f).vhat)), f


def test_grouped_launches_match_separate(golden):
    """The step with independent GEMMs sharing launches equals the step with
    one launch per GEMM (kernel variants may differ: 1e-13)."""
    import torch
    from pypde_amd.engine import RBCBatch, _GemmGroup, _SplitRhs
    pre = "step_rbc_eu_in_"
    engs = [RBCBatch((20, 16), [dict(ra=1e5, pr=1.0, dt=0.01)], case="rbc", group_launches=g)
            for g in (True, False)]
    assert any(isinstance(l, (_GemmGroup, _SplitRhs)) for l in engs[0].launches)
    assert not any(isinstance(l, (_GemmGroup, _SplitRhs)) for l in engs[1].launches)
    for e in engs:
        e.X[0, 0].copy_(torch.as_tensor(golden[pre + "U"]))
        e.X[0, 1].copy_(torch.as_tensor(golden[pre + "V"]))
        e.X[0, 2].copy_(torch.as_tensor(golden[pre + "T"]))
        e.pres[0].copy_(torch.as_tensor(golden[pre + "pres"]))
        for _ in range(5):
            e.step_eager()
    for t in ("X", "pres"):
        a, b = (host(getattr(e, t)) for e in engs)
        assert rel(a, b) < 1e-13, t


@pytest.mark.parametrize("case", ["linear", "rbc"])
def test_c1_trajectory_1000_steps(golden, case):
    """Configuration c1: 64x64, Ra=1e5, Pr=1, dt=0.01, Euler, seeded noise 1e-3:
    Nusselt and kinetic-energy series agree with the reference to 1e-10 after
    1000 steps."""
    import torch
    from pypde_amd.navier import NavierStokes
    from oracle import pypde_oracle as o
    ns = NavierStokes(case=case, shape=(64, 64), ra=1e5, pr=1.0, dt=0.01, tsave=None,
                      dealias=True, integrator="eu", beta=1.0, aspect=1.0)
    ref = o.NavierStokes(case=case, shape=(64, 64), ra=1e5, pr=1.0, dt=0.01, tsave=None)
    o.seeded_initial_state(ref, seed=0)
    ns.set_state(T=ref.T.vhat)
    nsteps = 1000
    nu = np.zeros(nsteps)
    nuv = np.zeros(nsteps)
    ke = np.zeros(nsteps)
    for n in range(nsteps):
        ns.update()
        nu[n], nuv[n] = ns.eval_Nu()
        ke[n] = ns.eval_KE()
    torch.cuda.synchronize()
    e_nu = rel(nu, golden["c1_%s_nu_plate" % case])
    e_nuv = rel(nuv, golden["c1_%s_nu_vol" % case])
    e_ke = rel(ke, golden["c1_%s_ke" % case])
    print("c1 %s: rel err Nu %.2e Nuvol %.2e KE %.2e" % (case, e_nu, e_nuv, e_ke))
    assert e_nu < 1e-10 and e_nuv < 1e-10 and e_ke < 1e-10
    for f in ("T", "U", "V"):
        assert rel(host(getattr(ns, f).vhat), golden["c1_%s_final_%s" % (case, f)]) < 1e-9


def test_batched_instances_match_single(golden):
    """Instances of different Ra / dt stepped in one batch equal the same
    instances stepped alone (and the oracle)."""
    import torch
    from pypde_amd.engine import RBCBatch
    from oracle import pypde_oracle as o
    shape = (24, 20)
    params = [dict(ra=1e4, pr=1.0, dt=0.01), dict(ra=3e5, pr=1.0, dt=0.004),
              dict(ra=2e6, pr=0.7, dt=0.002)]
    eng = RBCBatch(shape, params, case="rbc")
    refs = []
    for b, p in enumerate(params):
        r = o.NavierStokes(case="rbc", shape=shape, ra=p["ra"], pr=p["pr"], dt=p["dt"], tsave=None)
        o.seeded_initial_state(r, seed=10 + b, amplitude=0.05)
        eng.X[b, 2].copy_(torch.as_tensor(r.T.vhat))
        refs.append(r)
    for _ in range(4):
        eng.step_eager()
        for r in refs:
            r.update()
    for b, r in enumerate(refs):
        for f, idx in (("U", 0), ("V", 1), ("T", 2)):
            e = rel(host(eng.X[b, idx]), getattr(r, f).vhat)
            assert e < 1e-11, (b, f, e)
        assert rel(host(eng.pres[b]), r.pres.vhat) < 1e-11


@pytest.mark.parametrize("kw", [dict(integrator="rk3"), dict(beta=0.5)])
def test_op_by_op_fallback_matches_oracle(kw):
    """RK3 and beta != 1 with an odd dealiased y length (N = 22: L = 33) take
    the op-by-op path; it follows the oracle to 1e-12 over three steps."""
    from pypde_amd.navier import NavierStokes
    from oracle import pypde_oracle as o
    shape = (20, 22)
    ns = NavierStokes(case="rbc", shape=shape, ra=1e5, pr=1.0, dt=0.01, tsave=None, **kw)
    assert not ns.fused
    ref = o.NavierStokes(case="rbc", shape=shape, ra=1e5, pr=1.0, dt=0.01, tsave=None, **kw)
    o.seeded_initial_state(ref, seed=12, amplitude=0.1)
    ns.set_state(T=ref.T.vhat)
    for _ in range(3):
        ns.update()
        ref.update()
    for f in ("T", "U", "V", "pres"):
        e = rel(host(getattr(ns, f).vhat), getattr(ref, f).vhat)
        assert e < 1e-12, (kw, f, e)
