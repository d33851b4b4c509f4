"""GPU adjoint-descent iterations (drop-in navier.rbc2d_adj.NavierStokesAdjoint)
against the reference's own step (tests/golden) and several steps of the
oracle's restatement (Euler and RK3)."""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


def host(t):
    return t.detach().cpu().numpy()


def make(**kw):
    from pypde_amd.navier import NavierStokesAdjoint
    args = dict(case="rbc", shape=(20, 16), ra=5e4, pr=1.0, dt=0.005, tsave=None, dealias=True,
                integrator="eu", beta=1.0)
    args.update(kw)
    return NavierStokesAdjoint(**args)


def test_step_matches_reference(golden):
    """One update from the reference's input: 1e-12 (norm-wise) for T, U, V, the
    adjoint pressure and the inner solver's pressure."""
    from pypde_amd import ops
    nsa = make()
    assert nsa.NS.fused
    for f in ("T", "U", "V"):
        getattr(nsa, f).vhat = ops.to_dev(golden["adj_in_" + f])
    nsa.update()
    for f in ("T", "U", "V"):
        e = rel(host(getattr(nsa, f).vhat), golden["adj_out_" + f])
        assert e < 1e-12, (f, e)
    assert rel(host(nsa.pres.vhat), golden["adj_out_pres"]) < 1e-12
    assert rel(host(nsa.NS.pres.vhat), golden["adj_ns_pres"]) < 1e-12


@pytest.mark.parametrize("integrator", ["eu", "rk3"])
def test_steps_match_oracle(integrator):
    """10 updates from a convecting state (the 20x16, Ra=1e4 state of the
    steady-state fixture after 300 time steps, with its pressure in the inner
    solver): within 1e-10 of the oracle."""
    import os
    from oracle import pypde_oracle as o
    from pypde_amd import ops
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "rbc2d_steady.npz"))
    kw = dict(case="rbc", shape=(20, 16), ra=1e4, pr=1.0, dt=0.01, tsave=None, dealias=True,
              integrator=integrator, beta=1.0)
    ref = o.NavierStokesAdjoint(**kw)
    nsa = make(**kw)
    n = ref.T.vhat.size
    X = g["steady_X0"]
    for k, f in enumerate("TUV"):
        a = X[k * n:(k + 1) * n].reshape(ref.T.vhat.shape)
        getattr(ref, f).vhat = a.copy()
        getattr(nsa, f).vhat = ops.to_dev(a)
    ref.NS.pres.vhat = g["steady_pres0"].copy()
    nsa.NS.set_state(pres=g["steady_pres0"])
    for _ in range(10):
        ref.update()
        nsa.update()
    for f in ("T", "U", "V", "pres"):
        e = rel(host(getattr(nsa, f).vhat), getattr(ref, f).vhat)
        assert e < 1e-10, (f, e)
    for f in ("T", "U", "V"):
        e = rel(host(getattr(nsa.NS, f).vhat), getattr(ref.NS, f).vhat)
        assert e < 1e-10, ("residual " + f, e)


@pytest.mark.parametrize("integrator", ["eu", "rk3"])
def test_graph_replay_matches_op_by_op(integrator):
    """update() replayed from its HIP graph (from the second call on) gives the
    op-by-op results bit for bit, also after the caller replaces a state tensor
    between updates."""
    import os
    from pypde_amd import ops
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "rbc2d_steady.npz"))
    kw = dict(case="rbc", shape=(20, 16), ra=1e4, pr=1.0, dt=0.01, tsave=None, dealias=True,
              integrator=integrator, beta=1.0)
    runs = [make(use_graph=True, **kw), make(use_graph=False, **kw)]
    n = runs[0].T.vhat.numel()
    X = g["steady_X0"]
    shape = tuple(runs[0].T.vhat.shape)
    for nsa in runs:
        for k, f in enumerate("TUV"):
            getattr(nsa, f).vhat = ops.to_dev(X[k * n:(k + 1) * n].reshape(shape))
        nsa.NS.set_state(pres=g["steady_pres0"])
    for step in range(6):
        if step == 4:
            bump = 1e-3 * np.cos(np.arange(n)).reshape(shape)
            for nsa in runs:
                nsa.U.vhat = nsa.U.vhat + ops.to_dev(bump)
        for nsa in runs:
            nsa.update()
        for f in ("T", "U", "V", "pres", "UA", "VA", "TA"):
            a, b = host(getattr(runs[0], f).vhat), host(getattr(runs[1], f).vhat)
            assert np.array_equal(a, b), (integrator, step, f, np.abs(a - b).max())
        for f in ("T", "U", "V"):
            assert np.array_equal(host(getattr(runs[0].NS, f).vhat),
                                  host(getattr(runs[1].NS, f).vhat)), (integrator, step, f)
    assert runs[0]._graph is not None and runs[1]._graph is None
