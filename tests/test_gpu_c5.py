"""Configuration c5 at its own size: gradients with respect to the initial
temperature coefficients at 128x128, Ra=5e4, Pr=1, dt=0.005 (rbc case, seeded
noise 1e-3 from seed 0), by the hand-written adjoint of the fused step:

* J_nu = time-averaged plate Nusselt number over 2000 steps, checkpoints every
  100 steps with recomputed segments, through the torch.autograd.Function
  TimeAveragedNusselt; against central differences with eps = 1e-6 along 8
  unit directions drawn from numpy default_rng(5), 1e-6 relative to each
  directional derivative; the eps = 1e-4 quotient, whose O(eps^2) truncation
  error is a few 1e-6 here, within 1e-5;
* J_res = 0.5 ||(Phi_dt(X) - X)/dt||^2 at a state 400 steps into the run
  (rbc2d_base.py:321-341 restated, pressure held fixed): its temperature
  gradient against central differences along the same kind of directions,
  and its full gradient against the update direction of the reference's
  adjoint-descent iteration (navier/rbc2d_adj.py:321-356, here the drop-in
  NavierStokesAdjoint, which matches the reference's update to 1e-12).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 128
RA, DT = 5e4, 0.005


@pytest.fixture(scope="module")
def c5():
    import torch
    from pypde_amd.adjoint import Trajectory
    from pypde_amd.bases import Base
    from pypde_amd.engine import RBCBatch
    from pypde_amd.field import Field
    eng = RBCBatch((N, N), [dict(ra=RA, pr=1.0, dt=DT)], case="rbc")
    Tf = Field([Base(N, "CN"), Base(N, "CD")])
    X0 = torch.zeros_like(eng.X)
    noise = np.random.default_rng(0).uniform(-1e-3, 1e-3, size=(N, N))
    X0[0, 2].copy_(Tf.forward(noise))
    pres0 = torch.zeros_like(eng.pres)
    traj = Trajectory(eng, ckpt=100)
    return eng, X0, pres0, traj


def _directions(shape, n, seed, dev):
    import torch
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        d = rng.standard_normal(shape)
        out.append(torch.as_tensor(d / np.linalg.norm(d), device=dev))
    return out


def test_c5_nusselt_gradient_2000_steps_fd(c5):
    import torch
    from pypde_amd.adjoint import TimeAveragedNusselt
    eng, X0, pres0, traj = c5
    nsteps = 2000
    T0 = X0[:, 2].clone().requires_grad_(True)
    J = TimeAveragedNusselt.apply(T0, traj, X0, pres0, nsteps)
    J.sum().backward()
    grad = T0.grad.detach()
    print("J_nu = %.15f, |grad| = %.6e" % (J.item(), torch.linalg.norm(grad).item()))

    def fd(d, eps):
        Xp = X0.clone()
        Xp[:, 2] += eps * d
        Xm = X0.clone()
        Xm[:, 2] -= eps * d
        Jp, _ = traj.run(Xp, pres0, nsteps)
        Jm, _ = traj.run(Xm, pres0, nsteps)
        return ((Jp - Jm) / (2 * eps)).item()

    worst6 = worst4 = 0.0
    for k, d in enumerate(_directions(tuple(T0.shape), 8, 5, T0.device)):
        an = (grad * d).sum().item()
        f6, f4 = fd(d, 1e-6), fd(d, 1e-4)
        e6, e4 = abs(f6 - an) / abs(an), abs(f4 - an) / abs(an)
        worst6, worst4 = max(worst6, e6), max(worst4, e4)
        print("direction %d: adjoint %.12e fd(1e-6) %.12e (rel %.2e) fd(1e-4) %.12e (rel %.2e)"
              % (k, an, f6, e6, f4, e4))
    assert worst6 <= 1e-6, worst6
    assert worst4 <= 1e-5, worst4


def test_c5_residual_gradient_fd_and_adjoint_descent(c5):
    import torch
    from pypde_amd.adjoint import steady_residual, steady_residual_grad
    from pypde_amd.navier import NavierStokesAdjoint
    eng, X0, pres0, traj = c5
    eng.X.copy_(X0)
    eng.pres.copy_(pres0)
    for _ in range(400):
        eng.step_eager()
    Xs, ps = eng.X.clone(), eng.pres.clone()
    J, g = steady_residual_grad(eng, Xs, ps)
    gT = g[:, 2]
    print("J_res = %.12e, |dJ/dT| = %.6e" % (J.item(), torch.linalg.norm(gT).item()))
    worst = 0.0
    for k, d in enumerate(_directions(tuple(gT.shape), 8, 5, gT.device)):
        an = (gT * d).sum().item()
        Xp, Xm = Xs.clone(), Xs.clone()
        Xp[:, 2] += 1e-6 * d
        Xm[:, 2] -= 1e-6 * d
        f6 = ((steady_residual(eng, Xp, ps) - steady_residual(eng, Xm, ps)) / 2e-6).item()
        e = abs(f6 - an) / abs(an)
        worst = max(worst, e)
        print("direction %d: adjoint %.12e fd(1e-6) %.12e (rel %.2e)" % (k, an, f6, e))
    assert worst <= 1e-6, worst
    # the reference's adjoint-descent update from the same velocities and
    # temperature (its inner solver starts from zero pressure)
    J0, g0 = steady_residual_grad(eng, Xs, torch.zeros_like(ps))
    nsa = NavierStokesAdjoint(case="rbc", shape=(N, N), ra=RA, pr=1.0, dt=DT, tsave=None,
                              use_graph=False)
    nsa.U.vhat, nsa.V.vhat, nsa.T.vhat = (Xs[0, i].clone() for i in range(3))
    nsa.update()
    step = torch.cat([(getattr(nsa, f).vhat - Xs[0, i]).reshape(-1)
                      for i, f in enumerate("UVT")])
    gv = g0[0].reshape(-1)
    cos = (step @ gv / (torch.linalg.norm(step) * torch.linalg.norm(gv))).item()
    print("cosine(adjoint-descent update, discrete gradient of J_res) = %.4f" % cos)
    assert cos < 0.0
