"""GPU checks of the hand-written adjoint of the fused RBC2D step.

* one step: the step is quadratic in its state, so central differences are
  exact up to rounding -> the VJP matches them to 1e-9 relative;
* time-averaged Nusselt number over a checkpointed trajectory: gradient with
  respect to the initial temperature against central differences (eps=1e-6)
  along seeded random directions, 1e-6 relative;
* autograd wrappers, and the direction of the steady-state-residual gradient
  against the reference's adjoint-descent step (navier/rbc2d_adj.py:321-356).
"""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


def make(shape=(24, 20), ra=5e4, dt=0.005, B=1, case="rbc", integrator="eu", beta=1.0):
    import torch
    from pypde_amd.engine import RBCBatch
    params = [dict(ra=ra * (1 + 0.5 * b), pr=1.0, dt=dt) for b in range(B)]
    eng = RBCBatch(shape, params, case=case, integrator=integrator, beta=beta)
    rng = np.random.default_rng(3)
    X = torch.as_tensor(rng.standard_normal(tuple(eng.X.shape)) * 1e-2, device=eng.X.device)
    pres = torch.as_tensor(rng.standard_normal(tuple(eng.pres.shape)) * 1e-3, device=eng.X.device)
    return eng, X, pres


def run_step(eng, X, pres):
    eng.X.copy_(X)
    eng.pres.copy_(pres)
    eng.step_eager()
    return eng.X.clone(), eng.pres.clone()


@pytest.mark.parametrize("case", ["rbc", "linear"])
def test_step_vjp_matches_central_differences(case):
    import torch
    from pypde_amd.adjoint import StepAdjoint
    eng, X, pres = make(case=case, B=2)
    adj = StepAdjoint(eng)
    _check_vjp_fd(eng, adj, X, pres, eps=1e-3, tol=1e-9)


@pytest.mark.parametrize("integrator,beta", [("rk3", 1.0), ("eu", 0.5), ("rk3", 0.5)])
def test_step_vjp_rk3_and_explicit_diffusion(integrator, beta):
    """The three-stage step (a composition of three quadratic stages, the
    c-term coupling each stage to the one before) and the Euler step with an
    explicit share of the diffusion: VJP against central differences.  The Euler step is quadratic, so one quotient (eps = 1e-3) is
    exact up to rounding; the three-stage step has cubic and higher terms, so
    the quotients at eps and eps/2 are Richardson-extrapolated, which removes
    the eps^2 term of the truncation error (about 5e-7 relative at eps = 1e-4
    here), and judged at 1e-8 (the extrapolation also amplifies the rounding
    of the quotients, about 5e-9 on the smallest directional derivative)."""
    from pypde_amd.adjoint import StepAdjoint
    eng, X, pres = make(case="rbc", B=2, integrator=integrator, beta=beta)
    adj = StepAdjoint(eng)
    if integrator == "eu":
        _check_vjp_fd(eng, adj, X, pres, eps=1e-3, tol=1e-9)
    else:
        _check_vjp_fd(eng, adj, X, pres, eps=1e-4, tol=1e-8, richardson=True)


def _check_vjp_fd(eng, adj, X, pres, eps, tol, richardson=False):
    import torch
    rng = np.random.default_rng(11)
    dev = X.device
    wX = torch.as_tensor(rng.standard_normal(tuple(X.shape)), device=dev)
    wp = torch.as_tensor(rng.standard_normal(tuple(pres.shape)), device=dev)
    gX, gp = adj.vjp(X, pres, wX, wp)
    for trial in range(3):
        dX = torch.as_tensor(rng.standard_normal(tuple(X.shape)), device=dev)
        dp = torch.as_tensor(rng.standard_normal(tuple(pres.shape)), device=dev)
        def quotient(h):
            Xa, pa = run_step(eng, X + h * dX, pres + h * dp)
            Xm, pm = run_step(eng, X - h * dX, pres - h * dp)
            return ((wX * (Xa - Xm)).sum() + (wp * (pa - pm)).sum()).item() / (2 * h)
        fd = quotient(eps)
        if richardson:
            fd = (4.0 * quotient(eps / 2) - fd) / 3.0
        an = ((gX * dX).sum() + (gp * dp).sum()).item()
        print("trial %d: fd %.15e vjp %.15e rel %.2e" % (trial, fd, an, abs(fd - an) / abs(fd)))
        assert abs(fd - an) <= tol * max(abs(fd), 1e-300), (trial, fd, an)


def test_autograd_step_matches_vjp():
    import torch
    from pypde_amd.adjoint import RBCStep, StepAdjoint
    eng, X, pres = make()
    adj = StepAdjoint(eng)
    Xr = X.clone().requires_grad_(True)
    pr = pres.clone().requires_grad_(True)
    X1, p1 = RBCStep.apply(Xr, pr, adj)
    X2, p2 = RBCStep.apply(X1, p1, adj)
    loss = (X2 * X2).sum() + p2.sum()
    loss.backward()
    # manual reverse sweep through the same two steps
    Xa, pa = run_step(eng, X, pres)
    Xb2, pb2 = run_step(eng, Xa, pa)
    gX, gp = adj.vjp(Xa, pa, 2 * Xb2, torch.ones_like(pb2))
    gX, gp = adj.vjp(X, pres, gX, gp)
    assert rel(Xr.grad.cpu().numpy(), gX.cpu().numpy()) < 1e-13
    assert rel(pr.grad.cpu().numpy(), gp.cpu().numpy()) < 1e-13


def test_time_averaged_nusselt_gradient_fd():
    """J = mean_k Nu_plate(T_k) over 60 steps (checkpoints every 20) at 24x20,
    Ra=5e4, dt=0.005; dJ/dT0 against central differences along 4 seeded random
    unit directions d:
      eps = 1e-6: |fd - <grad, d>| <= 1e-6 ||grad||  (the difference quotient is
                  rounding-limited at about 1e-8 absolute, so the error is
                  measured against the gradient norm);
      eps = 1e-4: |fd - <grad, d>| <= 1e-6 |<grad, d>|  (converged quotient)."""
    import torch
    from pypde_amd.adjoint import TimeAveragedNusselt, Trajectory
    from pypde_amd.engine import RBCBatch
    eng = RBCBatch((24, 20), [dict(ra=5e4, pr=1.0, dt=0.005)], case="rbc")
    dev = eng.X.device
    rng = np.random.default_rng(5)
    X0 = torch.zeros_like(eng.X)
    X0[:, 2] = torch.as_tensor(rng.uniform(-0.05, 0.05, tuple(X0[:, 2].shape)), device=dev)
    pres0 = torch.zeros_like(eng.pres)
    traj = Trajectory(eng, ckpt=20)
    nsteps = 60
    T0 = X0[:, 2].clone().requires_grad_(True)
    J = TimeAveragedNusselt.apply(T0, traj, X0, pres0, nsteps)
    J.sum().backward()
    grad = T0.grad.detach()
    gnorm = torch.linalg.norm(grad).item()

    def fd(d, eps):
        Xp = X0.clone(); Xp[:, 2] += eps * d
        Xm = X0.clone(); Xm[:, 2] -= eps * d
        Jp, _ = traj.run(Xp, pres0, nsteps)
        Jm, _ = traj.run(Xm, pres0, nsteps)
        return ((Jp - Jm) / (2 * eps)).item()

    for trial in range(4):
        d = torch.as_tensor(rng.standard_normal(tuple(T0.shape)), device=dev)
        d = d / torch.linalg.norm(d)
        an = (grad * d).sum().item()
        f6, f4 = fd(d, 1e-6), fd(d, 1e-4)
        print("direction %d: adjoint %.12e fd(1e-6) %.12e fd(1e-4) %.12e |grad| %.3e" %
              (trial, an, f6, f4, gnorm))
        assert abs(f6 - an) <= 1e-6 * gnorm
        assert abs(f4 - an) <= 1e-6 * abs(an)


def test_nusselt_functional_matches_eval_nu():
    import torch
    from pypde_amd.adjoint import NusseltFunctional
    from pypde_amd.navier import NavierStokes
    ns = NavierStokes(case="rbc", shape=(24, 20), ra=5e4, pr=1.0, dt=0.005, tsave=None)
    rng = np.random.default_rng(8)
    ns.set_state(T=rng.standard_normal((22, 18)) * 0.01)
    fn = NusseltFunctional(ns.engine)
    a = fn.value(ns.engine.X[:, 2]).item()
    b = ns.eval_Nu()[0]
    assert abs(a - b) < 1e-12 * abs(b)


def test_residual_gradient_direction_vs_reference_adjoint(golden):
    """The reference's adjoint-descent step (rbc2d_adj.py:321-356) moves the
    state along a smoothed, continuous-adjoint descent direction of the
    steady-state residual; the discrete gradient of J_res at the same state
    must point the same way (negative inner product with the step)."""
    import torch
    from pypde_amd.adjoint import steady_residual_grad
    from pypde_amd.engine import RBCBatch
    eng = RBCBatch((20, 16), [dict(ra=5e4, pr=1.0, dt=0.005)], case="rbc")
    dev = eng.X.device
    X = torch.zeros_like(eng.X)
    X[0, 0] = torch.as_tensor(golden["adj_in_U"], device=dev)
    X[0, 1] = torch.as_tensor(golden["adj_in_V"], device=dev)
    X[0, 2] = torch.as_tensor(golden["adj_in_T"], device=dev)
    J, g = steady_residual_grad(eng, X, torch.zeros_like(eng.pres))
    step = np.concatenate([(golden["adj_out_" + f] - golden["adj_in_" + f]).ravel() for f in "UVT"])
    gv = np.concatenate([g[0, i].cpu().numpy().ravel() for i in range(3)])
    cos = float(step @ gv / (np.linalg.norm(step) * np.linalg.norm(gv)))
    print("cosine(reference adjoint-descent step, discrete gradient) = %.4f" % cos)
    assert cos < 0.0


def test_graph_replayed_sweep_is_bit_identical():
    """Trajectory with graph-replayed forward steps and VJPs gives the op-by-op
    sweep's J and gradients bit for bit (same launches in the same order), also
    on a second call that reuses the captured graphs, and with every state kept
    instead of checkpoints."""
    import torch
    from pypde_amd.adjoint import Trajectory
    from pypde_amd.engine import RBCBatch
    out = []
    for use_graph in (True, False):
        eng = RBCBatch((24, 20), [dict(ra=5e4, pr=1.0, dt=0.005),
                                  dict(ra=2e4, pr=1.0, dt=0.005)], case="rbc")
        rng = np.random.default_rng(12)
        X0 = torch.zeros_like(eng.X)
        X0[:, 2] = torch.as_tensor(rng.uniform(-0.05, 0.05, tuple(X0[:, 2].shape)),
                                   device=eng.X.device)
        pres0 = torch.zeros_like(eng.pres)
        traj = Trajectory(eng, ckpt=7, use_graph=use_graph)
        runs = []
        for nsteps in (23, 15):
            J, Xb, pb = traj.grad_nusselt(X0, pres0, nsteps)
            runs.append([t.detach().cpu().numpy() for t in (J, Xb, pb)])
        out.append(runs)
        assert (traj.adj.graph is not None) == use_graph
    for a, b in zip(out[0], out[1]):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    # every state kept (the default when the trajectory fits in device memory):
    # no recomputation, the same gradient
    eng = RBCBatch((24, 20), [dict(ra=5e4, pr=1.0, dt=0.005),
                              dict(ra=2e4, pr=1.0, dt=0.005)], case="rbc")
    rng = np.random.default_rng(12)
    X0 = torch.zeros_like(eng.X)
    X0[:, 2] = torch.as_tensor(rng.uniform(-0.05, 0.05, tuple(X0[:, 2].shape)),
                               device=eng.X.device)
    traj = Trajectory(eng)
    assert traj._spacing(23) == 1
    J, Xb, pb = traj.grad_nusselt(X0, torch.zeros_like(eng.pres), 23)
    for x, y in zip([t.detach().cpu().numpy() for t in (J, Xb, pb)], out[0][0]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("n", [22, 126, 254, 510])
@pytest.mark.parametrize("variant", [1, 2, 3, 4])
def test_reversed_transposed_solves_match_dense(n, variant):
    """The adjoint's transposed line solves (negative row stride, the line read
    from its last row) in each kernel form, with per-entry Helmholtz tables and
    with per-line Poisson tables, against dense A^T solves on the host, at the
    line lengths of 24^2, 128^2, 256^2 and 512^2 steps."""
    import torch
    from pypde_amd import ops
    from pypde_amd.adjoint import StepAdjoint
    from pypde_amd.bases import Base
    from pypde_amd.solver import rec_poisson_reversed_transpose, rec_reversed_transpose
    from pypde_amd.templates import hholtz_matrices, poisson_matrices
    dev = ops.device()
    rng = np.random.default_rng(n + 7 * variant)
    # per-entry tables: three (instance, field) entries with their own matrices
    mats = [np.asarray(hholtz_matrices(Base(n + 2, key), lam, 0.5)[2])
            for key, lam in (("CD", 1e-3), ("CN", 2e-2), ("CD", 0.3))]
    tabs = torch.as_tensor(np.stack([np.stack(rec_reversed_transpose(A)) for A in mats]),
                           dtype=torch.float64, device=dev)
    nlines = 77
    b = rng.standard_normal((3, n, nlines))
    x = torch.as_tensor(b, device=dev).contiguous()
    StepAdjoint._reversed_solve(x, n, nlines, 3, tabs, False, variant=variant)
    xh = x.cpu().numpy()
    for e, A in enumerate(mats):
        ref = np.linalg.solve(A.T, b[e])
        err = np.abs(xh[e] - ref).max() / np.abs(ref).max()
        assert err < 1e-12, ("entry", n, variant, e, err)
    # per-line tables: the Poisson eigen-split lines, one singular line
    pm = poisson_matrices([Base(n + 2, "CN"), Base(n + 2, "CN")], singular=True,
                          scale=(0.5, 0.5))
    A, C, w = (np.asarray(pm[k]) for k in ("Ax", "Cx", "wy"))
    pt = torch.as_tensor(np.stack(rec_poisson_reversed_transpose(A, C, w, True)), device=dev)
    tabs2 = (pt, pt.transpose(1, 2).contiguous())
    nl = w.shape[0]
    b = rng.standard_normal((1, n, nl))
    x = torch.as_tensor(b, device=dev).contiguous()
    StepAdjoint._reversed_solve(x, n, nl, 1, tabs2, True, variant=variant)
    xh = x.cpu().numpy()[0]
    worst = 0.0
    for j in range(nl):
        M = A + w[j] * C
        if abs(w[j]) < 1e-10:
            ref = np.zeros(n)
            ref[1:] = np.linalg.solve(M[1:, 1:].T, b[0, 1:, j])
        else:
            ref = np.linalg.solve(M.T, b[0, :, j])
        worst = max(worst, np.abs(xh[:, j] - ref).max() / np.abs(ref).max())
    assert worst < 1e-11, ("per line", n, variant, worst)


def test_step_vjp_128_matches_central_differences():
    """One step at 128x128 (the c5 size): the step is quadratic in its state,
    so the central difference is exact up to rounding; the VJP against it along
    3 random directions, 1e-9 relative."""
    import torch
    from pypde_amd.adjoint import StepAdjoint
    eng, X, pres = make(shape=(128, 128), ra=5e4, dt=0.005)
    adj = StepAdjoint(eng)
    rng = np.random.default_rng(21)
    dev = X.device
    wX = torch.as_tensor(rng.standard_normal(tuple(X.shape)), device=dev)
    wp = torch.as_tensor(rng.standard_normal(tuple(pres.shape)), device=dev)
    gX, gp = adj.vjp(X, pres, wX, wp)
    for trial in range(3):
        dX = torch.as_tensor(rng.standard_normal(tuple(X.shape)), device=dev)
        dp = torch.as_tensor(rng.standard_normal(tuple(pres.shape)), device=dev)
        eps = 1e-3
        Xa, pa = run_step(eng, X + eps * dX, pres + eps * dp)
        Xm, pm = run_step(eng, X - eps * dX, pres - eps * dp)
        fd = ((wX * (Xa - Xm)).sum() + (wp * (pa - pm)).sum()).item() / (2 * eps)
        an = ((gX * dX).sum() + (gp * dp).sum()).item()
        print("128x128 step VJP, trial %d: fd %.15e vjp %.15e" % (trial, fd, an))
        assert abs(fd - an) <= 1e-9 * abs(fd), (trial, fd, an)


@pytest.mark.parametrize("integrator,beta", [("rk3", 1.0), ("rk3", 0.5)])
def test_time_averaged_nusselt_gradient_rk3(integrator, beta):
    """J = mean_k Nu_plate(T_k) over 40 three-stage steps (checkpoints every
    10) at 24x20: the checkpointed reverse sweep against central differences
    (eps = 1e-5) along 3 unit directions, 1e-6 relative to each directional
    derivative."""
    import torch
    from pypde_amd.adjoint import TimeAveragedNusselt, Trajectory
    from pypde_amd.engine import RBCBatch
    eng = RBCBatch((24, 20), [dict(ra=5e4, pr=1.0, dt=0.005)], case="rbc",
                   integrator=integrator, beta=beta)
    dev = eng.X.device
    rng = np.random.default_rng(15)
    X0 = torch.zeros_like(eng.X)
    X0[:, 2] = torch.as_tensor(rng.uniform(-0.05, 0.05, tuple(X0[:, 2].shape)), device=dev)
    pres0 = torch.zeros_like(eng.pres)
    traj = Trajectory(eng, ckpt=10)
    nsteps = 40
    T0 = X0[:, 2].clone().requires_grad_(True)
    J = TimeAveragedNusselt.apply(T0, traj, X0, pres0, nsteps)
    J.sum().backward()
    grad = T0.grad.detach()
    for trial in range(3):
        d = torch.as_tensor(rng.standard_normal(tuple(T0.shape)), device=dev)
        d = d / torch.linalg.norm(d)
        an = (grad * d).sum().item()
        eps = 1e-5
        Xp = X0.clone()
        Xp[:, 2] += eps * d
        Xm = X0.clone()
        Xm[:, 2] -= eps * d
        Jp, _ = traj.run(Xp, pres0, nsteps)
        Jm, _ = traj.run(Xm, pres0, nsteps)
        fd = ((Jp - Jm) / (2 * eps)).item()
        print("%s beta %.1f direction %d: adjoint %.12e fd %.12e rel %.2e"
              % (integrator, beta, trial, an, fd, abs(fd - an) / abs(an)))
        assert abs(fd - an) <= 1e-6 * abs(an)
