"""Pin the oracle's RBC2D step against vectors written from the reference
itself (tests/golden/make_golden.py): these are the functions no closed form
pins (NavierStokes.update and its trajectories)."""
import numpy as np
import pytest

from oracle import pypde_oracle as o
from conftest import rel

CASES = {
    "linear_eu": dict(case="linear", integrator="eu", beta=1.0),
    "rbc_eu": dict(case="rbc", integrator="eu", beta=1.0),
    "rbc_rk3": dict(case="rbc", integrator="rk3", beta=1.0),
    "rbc_eu_beta": dict(case="rbc", integrator="eu", beta=0.5),
    "zero_eu": dict(case="zero", integrator="eu", beta=1.0),
}


def load_state(ns, golden, prefix):
    ns.T.vhat = golden[prefix + "T"].copy()
    ns.U.vhat = golden[prefix + "U"].copy()
    ns.V.vhat = golden[prefix + "V"].copy()
    ns.P.vhat = golden[prefix + "P"].copy()
    ns.pres.vhat = golden[prefix + "pres"].copy()


@pytest.mark.parametrize("name", sorted(CASES))
def test_single_steps_match_reference(golden, name):
    kw = dict(CASES[name])
    case = kw.pop("case")
    ns = o.NavierStokes(case=case, shape=(20, 16), ra=1e5, pr=1.0, dt=0.01, tsave=None,
                        dealias=True, aspect=1.0, **kw)
    load_state(ns, golden, "step_%s_in_" % name)
    for n in range(3):
        ns.update()
        if n == 0:
            for f in ("T", "U", "V", "P", "pres"):
                assert rel(getattr(ns, f).vhat, golden["step_%s_out1_%s" % (name, f)]) < 1e-14, f
    for f in ("T", "U", "V", "P", "pres"):
        assert rel(getattr(ns, f).vhat, golden["step_%s_out3_%s" % (name, f)]) < 1e-14, f


@pytest.mark.parametrize("case", ["linear", "rbc"])
def test_c1_trajectory_prefix_matches_reference(golden, case):
    """First 200 steps of the 64x64, Ra=1e5 run (the full 1000 steps run in the
    GPU tests against the same fixture)."""
    ns = o.NavierStokes(case=case, shape=(64, 64), ra=1e5, pr=1.0, dt=0.01, tsave=None,
                        dealias=True, integrator="eu", beta=1.0, aspect=1.0)
    o.seeded_initial_state(ns, seed=0)
    nsteps = 200
    nu = np.zeros(nsteps)
    ke = np.zeros(nsteps)
    for n in range(nsteps):
        ns.update()
        nu[n] = ns.eval_Nu()[0]
        ke[n] = ns.eval_KE()
    assert rel(nu, golden["c1_%s_nu_plate" % case][:nsteps]) < 1e-13
    assert rel(ke, golden["c1_%s_ke" % case][:nsteps]) < 1e-11


def test_steady_residual_matches_reference(golden):
    ns = o.NavierStokes(case="rbc", shape=(20, 16), ra=5e4, pr=1.0, dt=0.005, tsave=None)
    X = golden["steady_X"]
    mt, mu = ns.T.vhat.size, ns.U.vhat.size
    ns.T.vhat = X[:mt].reshape(ns.T.vhat.shape).copy()
    ns.U.vhat = X[mt:mt + mu].reshape(ns.U.vhat.shape).copy()
    ns.V.vhat = X[mt + mu:].reshape(ns.V.vhat.shape).copy()
    r = ns.steady_residual()
    assert rel(r, golden["steady_res"]) < 1e-13


def test_oracle_adjoint_descent_step_matches_reference(golden):
    """NavierStokesAdjoint.update (rbc2d_adj.py:321-356) from the golden input."""
    nsa = o.NavierStokesAdjoint(case="rbc", shape=(20, 16), ra=5e4, pr=1.0, dt=0.005,
                                tsave=None, dealias=True, integrator="eu", beta=1.0)
    nsa.T.vhat = golden["adj_in_T"].copy()
    nsa.U.vhat = golden["adj_in_U"].copy()
    nsa.V.vhat = golden["adj_in_V"].copy()
    nsa.update()
    for f in ("T", "U", "V"):
        assert rel(getattr(nsa, f).vhat, golden["adj_out_" + f]) < 1e-14, f
    assert rel(nsa.pres.vhat, golden["adj_out_pres"]) < 1e-14
    assert rel(nsa.NS.pres.vhat, golden["adj_ns_pres"]) < 1e-14


def test_dense_transform_oracle_matches_reference_single_steps(golden):
    """The oracle with dense-product transforms (the rounding-floor variant)
    still reproduces the reference's single steps: it is the same arithmetic
    up to rounding."""
    from oracle import pypde_oracle as o
    pre = "step_rbc_eu_"
    with o.dense_transforms():
        ns = o.NavierStokes(case="rbc", shape=(20, 16), ra=1e5, pr=1.0, dt=0.01, tsave=None)
        for f in ("T", "U", "V", "P", "pres"):
            getattr(ns, f).vhat = np.array(golden[pre + "in_" + f])
        ns.update()
    for f in ("T", "U", "V", "pres"):
        a, b = getattr(ns, f).vhat, golden[pre + "out1_" + f]
        assert np.linalg.norm(a - b) <= 1e-13 * np.linalg.norm(b), f


def test_c3_rounding_floor_fixture():
    """The c3 trajectory (512x512, Ra=1e8, 2000 steps) of the oracle with
    dense-product transforms (tests/golden/make_c3_floor.py) against the
    reference's own run: two exact FP64 restatements differ by this much, the
    floor on which the GPU series is judged.  The plate Nusselt number, a wall
    derivative that amplifies the rounding of the high modes by about N^2, is
    the most sensitive: up to about 1e-10 absolute throughout the run (4e-11
    over the first 10 steps), while Nu_vol and KE stay within 1e-11."""
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    a = np.load(os.path.join(here, "rbc2d_c3_dense_oracle.npz"))
    r = np.load(os.path.join(here, "rbc2d_c3_512.npz"))
    assert a["nu_plate"].shape == r["c3_nu_plate"].shape == (2000,)
    d = np.abs(a["nu_plate"] - r["c3_nu_plate"])
    dv = np.abs(a["nu_vol"] - r["c3_nu_vol"]) / np.abs(r["c3_nu_vol"])
    dk = np.abs(a["ke"] - r["c3_ke"]) / np.abs(r["c3_ke"])
    assert 5e-11 < d.max() < 2e-10
    assert dv[1:].max() < 1e-11 and dk[50:].max() < 1e-11
