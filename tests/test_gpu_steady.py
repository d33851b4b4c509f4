"""GPU steady states: steady_fun against the reference's values and
solve_steady_state against the reference's attractor (20x16, Ra=1e4,
dt=0.01; tests/golden/make_golden.py steady)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def steady():
    return np.load(os.path.join(HERE, "golden", "rbc2d_steady.npz"))


def host(t):
    return t.detach().cpu().numpy()


def make():
    from pypde_amd.navier import NavierStokes
    return NavierStokes(case="rbc", shape=(20, 16), ra=1e4, pr=1.0, dt=0.01, tsave=None,
                        dealias=True, integrator="eu", beta=1.0, aspect=1.0)


def test_steady_fun_matches_reference(steady):
    """F = (step(X) - X)/dt; one step parity is 1e-12 relative, so F agrees to
    1e-12 max|X| / dt (one step) and 1e-11 max|X| / 0.2 (20 steps)."""
    ns = make()
    assert ns.fused
    X0 = steady["steady_X0"]
    ns.set_state(pres=steady["steady_pres0"])
    F = host(ns.steady_fun(X0))
    assert np.abs(F - steady["steady_F0"]).max() < 1e-12 * np.abs(X0).max() / ns.dt
    ns.set_state(pres=steady["steady_pres0"])
    F = host(ns.steady_fun(X0, ns, 0.2))
    assert np.abs(F - steady["steady_F0_dt"]).max() < 1e-11 * np.abs(X0).max() / 0.2
    assert ns.time == 0.0


@pytest.mark.parametrize("dt", [None, 0.5])
def test_solve_steady_state_reaches_reference_attractor(steady, dt):
    """Newton-Krylov from the state after 300 steps lands on the state the
    reference's own map converges to after 14300 steps: [T, U, V] to 1e-9 of
    max|X|, the plate Nusselt number to 1e-9."""
    ns = make()
    X0 = steady["steady_X0"]
    ns.set_state(pres=steady["steady_pres0"])
    sol = ns.solve_steady_state(X0=X0, dt=dt, tol=1e-11, maxiter=50, disp=False)
    assert sol.success, (sol.message, sol.history)
    assert sol.fnorm <= 1e-11
    Xs = steady["steady_Xs"]
    x = host(sol.x)
    assert np.abs(x - Xs).max() < 1e-9 * np.abs(Xs).max()
    assert np.abs(host(sol.pres) - steady["steady_pres_s"]).max() < 1e-7
    # the solver leaves the state at the solution
    assert np.abs(host(ns.vectorify()) - x).max() == 0.0
    nu_plate, nu_vol = ns.eval_Nu()
    assert abs(float(nu_plate) - steady["steady_Nu"][0]) < 1e-9
    assert abs(float(nu_vol) - steady["steady_Nu"][1]) < 1e-9
    # a further step of the map does not move it
    ns.update()
    assert np.abs(host(ns.vectorify()) - x).max() < 1e-11 * ns.dt * 10


def test_reference_shaped_solve_keeps_reference_semantics(steady):
    """method="reference" is the reference's own call (scipy's root with
    method="krylov" on steady_fun over [T, U, V], the pressure advancing in
    the solver with every evaluation, rbc2d_base.py:265-291).  From the state
    after 300 steps it does not reach the attractor within 20 iterations --
    F keeps changing with the call history -- where the default Newton-Krylov
    on [T, U, V, pres] converges (previous test); its first evaluation is
    steady_fun at X0, the reference's value."""
    ns = make()
    X0 = steady["steady_X0"]
    ns.set_state(pres=steady["steady_pres0"])
    calls = []
    orig = ns.steady_fun

    def spy(X, NS=None, dt=None):
        F = orig(X, NS, dt)
        calls.append(host(F))
        return F
    ns.steady_fun = spy
    sol = ns.solve_steady_state(X0=X0, maxiter=20, tol=1e-11, disp=False, method="reference")
    assert np.abs(calls[0] - steady["steady_F0"]).max() < 1e-12 * np.abs(X0).max() / ns.dt
    assert len(calls) > 20
    assert not sol.success
    assert np.abs(sol.fun).max() > 1e-6
    with pytest.raises(ValueError):
        ns.solve_steady_state(X0=X0, method="lgmres")
