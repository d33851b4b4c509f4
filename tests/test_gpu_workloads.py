"""Parity at the benchmarked workloads themselves, with the kernels the
benchmark runs (the shapes are the autotuner's keys, so the same launch forms
are chosen as in bench.py):

* the microbench chain: 512 fields of 256x256 Chebyshev coefficients (seed 1,
  Gaussian times exp(-0.05 (k + l))) through the 12 kernels of bench.C2Chain;
  8 of the fields against the oracle's chain, norm-wise 1e-12 per field and
  per output (reference: chebyshev.py:67-93,287-337, plans.py:54-74,184-224);
* the 256x256 Euler step of the headline metric (rbc, Ra=1e8, dt=1e-3) with
  its tuned launch set and graph replays, against the oracle's update
  (navier/rbc2d.py:396-434) after 12 steps: 1e-12, or twice the distance
  between two exact FP64 restatements where that is larger;
* the mixed-size batch: the 256 instances of bench.c4_specs in one graph
  launch per step; one instance of every size class against the oracle after
  5 steps (1e-11, or twice the rounding floor), then 500 steps: finite states, plate Nusselt numbers near
  the conductive value, and sampled instances equal to single-instance runs.
"""
import os
import sys

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def host(t):
    return t.detach().cpu().numpy()


def _oracle_chain(c):
    """The chain of one field on the oracle: returns the outputs of K2, K4, K6,
    K8, K10 and K12."""
    from oracle import pypde_oracle as o
    N = c.shape[0]
    ch, cd = o.Basis(N, "CH"), o.Basis(N, "CD")
    Bm, _, Ax = o.hholtz_matrices(cd, 0.01, 1.0)
    along = o._along
    u1 = along(ch._cheb_backward, c, 0)
    u2 = along(ch._cheb_backward, u1, 1)
    c1 = along(ch._cheb_forward, u2, 0)
    c2 = along(ch._cheb_forward, c1, 1)
    g1 = along(cd.from_chebyshev, c2, 0)
    g2 = along(cd.from_chebyshev, g1, 1)
    t1 = along(cd.to_chebyshev, g2, 0)
    t2 = along(cd.to_chebyshev, t1, 1)
    r1 = o.MatAxis(Bm, 0)(c2)
    r2 = o.MatAxis(Bm, 1)(r1)
    x = o.FdmaAxis(Ax, 0)(r2)
    x = o.FdmaAxis(Ax, 1)(x)
    return dict(u2=u2, c2=c2, g2=g2, t2=t2, r2=r2, x=x)


def test_c2_chain_512_fields_matches_oracle():
    import torch
    import bench
    from pypde_amd import ops
    nf, N = 512, 256
    c_h = bench.c2_input(nf, N, seed=1)
    chain = bench.C2Chain(torch.as_tensor(c_h).to(ops.device()))
    chain.run_once()
    torch.cuda.synchronize()
    b = chain.bufs
    for f in np.linspace(0, nf - 1, 8).astype(int):
        ref = _oracle_chain(c_h[f])
        for key, want in ref.items():
            e = rel(host(b[key][f]), want)
            assert e < 1e-12, (int(f), key, e)


def test_plain_256_step_with_tuned_launches_matches_oracle():
    """Tolerance: 1e-12, or twice the rounding floor where that is larger.  At
    this size the temperature's Helmholtz operator (lambda = dt kappa = 1e-7)
    is close to the pseudo-inverse of D^2, so its solve amplifies rounding of
    the right-hand side by about N^2: two exact FP64 restatements of the
    reference (the oracle with scipy's DCT and with dense-product transforms)
    already differ by about 1e-10 in T after 12 steps from this state."""
    import torch
    import bench
    from pypde_amd.navier import NavierStokes
    from oracle import pypde_oracle as o
    from conftest import floor_tol, oracle_runs
    N = 256
    kw = dict(shape=(N, N), ra=1e8, pr=1.0, dt=1e-3, tsave=None)
    ns = NavierStokes(case="rbc", dealias=True, integrator="eu", beta=1.0, aspect=1.0, **kw)
    # the bench's initial state, plus a velocity so that the products matter
    noise = bench.seeded_noise((N, N), seed=0)
    ns.set_state(T=ns.T.forward(noise))
    ns.set_velocity(amplitude=0.05)
    init = {f: host(getattr(ns, f).vhat).copy() for f in "TUV"}

    def make():
        r = o.NavierStokes(case="rbc", **kw)
        for f in "TUV":
            getattr(r, f).vhat = init[f].copy()
        return r
    ns.prepare_graphs()
    print("launch variants:", ns.engine.variants())
    nsteps = ns.GRAPH_BLOCK + 2   # one 10-step graph replay and two 1-step replays
    ns.update_many(nsteps)
    torch.cuda.synchronize()
    ref, dense = oracle_runs(make, nsteps)
    for f in ("T", "U", "V", "pres"):
        e = rel(host(getattr(ns, f).vhat), getattr(ref, f).vhat)
        tol = floor_tol(ref, dense, f, 1e-12)
        print("256x256 %s after %d steps: %.2e (floor %.2e)"
              % (f, nsteps, e, rel(getattr(dense, f).vhat, getattr(ref, f).vhat)))
        assert e < tol, (f, e, tol)
    # Nusselt numbers: the plate value takes d/dy at the walls, which amplifies
    # the rounding of the high Chebyshev modes by about N^2; its floor is the
    # larger of the two restatements' distance and the distance between the two
    # evaluations of the reference's own final state
    nu, nr = ns.eval_Nu(), ref.eval_Nu()
    with o.dense_transforms():
        nd, nre = dense.eval_Nu(), ref.eval_Nu()
    for k in range(2):
        floor = max(abs(nd[k] - nr[k]), abs(nre[k] - nr[k]))
        tol = max(1e-10 * abs(nr[k]), 4.0 * floor)
        print("256x256 Nu[%d]: %.3e (floor %.3e)" % (k, abs(nu[k] - nr[k]), floor))
        assert abs(nu[k] - nr[k]) <= tol, (k, nu[k], nr[k], tol)


@pytest.fixture(scope="module")
def c4_batch():
    import bench
    from pypde_amd.batch import HeteroBatch
    specs = bench.c4_specs()
    hb = HeteroBatch(specs, case="rbc")
    seeds = [1000 + i for i in range(len(specs))]
    hb.seed(seeds, amplitude=1e-3)
    hb.capture(1)
    return specs, seeds, hb


def test_c4_one_instance_per_size_class_matches_oracle(c4_batch):
    """Tolerance: 1e-11, or twice the rounding floor of the instance where that
    is larger (see the 256x256 test)."""
    import torch
    from oracle import pypde_oracle as o
    from conftest import floor_tol, oracle_runs
    specs, seeds, hb = c4_batch
    assert len(hb.shapes) == 16
    first = {}
    for i, s in enumerate(specs):
        first.setdefault(tuple(s["shape"]), i)
    nsteps = 5
    hb.step(nsteps)
    torch.cuda.synchronize()
    for shape, i in sorted(first.items()):
        s = specs[i]

        def make(s=s, shape=shape, i=i):
            r = o.NavierStokes(case="rbc", shape=shape, ra=s["ra"], pr=s["pr"], dt=s["dt"],
                               tsave=None)
            return o.seeded_initial_state(r, seed=seeds[i], amplitude=1e-3)
        r, dense = oracle_runs(make, nsteps)
        st = hb.state(i)
        for f in ("U", "V", "T", "pres"):
            e = rel(host(st[f]), getattr(r, f).vhat)
            tol = floor_tol(r, dense, f, 1e-11)
            print("c4 %s instance %d %s: %.2e (tol %.2e)" % (shape, i, f, e, tol))
            assert e < tol, (shape, i, f, e, tol)


def test_c4_500_steps_properties(c4_batch):
    """Continues the batch of the previous test to 500 steps."""
    import torch
    from pypde_amd.engine import RBCBatch
    specs, seeds, hb = c4_batch
    # the previous test stepped 5 (or none when run alone)
    hb.seed(seeds, amplitude=1e-3)
    nsteps = 500
    hb.step(nsteps)
    torch.cuda.synchronize()
    assert hb.finite()
    nu = hb.nusselt()
    assert np.all(np.isfinite(nu))
    print("c4 Nu_plate after 500 steps: min %.6f max %.6f" % (nu.min(), nu.max()))
    assert np.all(nu > 0.9) and np.all(nu < 50.0)
    # sampled instances against the same instance stepped alone (its own
    # single-instance launch set and graph)
    sample = [0, 37, 101, 180, 255]
    for i in sample:
        s = specs[i]
        e = RBCBatch(tuple(s["shape"]), [dict(ra=s["ra"], pr=s["pr"], dt=s["dt"])], case="rbc")
        e.X[0, 2].copy_(hb.temperature_field(hb.where[i][0]).forward(
            np.random.default_rng(seeds[i]).uniform(-1e-3, 1e-3, size=s["shape"])))
        e.capture(1)
        e.step(nsteps)
        torch.cuda.synchronize()
        st = hb.state(i)
        for f, idx in (("U", 0), ("V", 1), ("T", 2)):
            err = rel(host(st[f]), host(e.X[0, idx]))
            assert err < 1e-9, (i, s["shape"], f, err)
