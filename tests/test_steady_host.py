"""Host logic of the steady-state solver (pypde_amd.steady) and the golden
steady state of the reference's time-step map (20x16, Ra=1e4, dt=0.01,
tests/golden/make_golden.py steady).

The map's fixed point pins two things: the oracle (its step must leave the
reference's converged state unchanged, and its steady_fun must reproduce the
reference's at the state after 300 steps), and the Newton-Krylov driver (run
on the oracle's map it must land on the reference's attractor)."""
import os

import numpy as np
import pytest
import torch

from oracle import pypde_oracle as o
from pypde_amd.steady import gmres, newton_krylov

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def steady():
    return np.load(os.path.join(HERE, "golden", "rbc2d_steady.npz"))


def oracle_ns():
    return o.NavierStokes(case="rbc", shape=(20, 16), ra=1e4, pr=1.0, dt=0.01, tsave=None,
                          dealias=True, integrator="eu", beta=1.0, aspect=1.0)


def load(ns, z):
    n = ns.T.vhat.size
    sh = ns.T.vhat.shape
    ns.T.vhat = z[:n].reshape(sh).copy()
    ns.U.vhat = z[n:2 * n].reshape(sh).copy()
    ns.V.vhat = z[2 * n:3 * n].reshape(sh).copy()
    if z.size > 3 * n:
        ns.pres.vhat = z[3 * n:].reshape(ns.pres.vhat.shape).copy()


def vec(ns, pres=False):
    parts = [ns.T.vhat.ravel(), ns.U.vhat.ravel(), ns.V.vhat.ravel()]
    if pres:
        parts.append(ns.pres.vhat.ravel())
    return np.concatenate(parts)


# ---- GMRES / Newton on problems with known answers -------------------------
@pytest.mark.parametrize("restart", [5, 40])
def test_gmres_matches_direct_solve(restart):
    rng = np.random.default_rng(3)
    n = 40
    A = np.eye(n) * 4 + rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    At = torch.from_numpy(A)
    x, its, relres = gmres(lambda v: At @ v, torch.from_numpy(b), tol=1e-13, restart=restart,
                           maxiter=200)
    xd = np.linalg.solve(A, b)
    assert relres <= 1e-13
    assert np.abs(x.numpy() - xd).max() <= 1e-11 * np.abs(xd).max()
    # the Arnoldi estimate is the true residual
    true = np.linalg.norm(A @ x.numpy() - b) / np.linalg.norm(b)
    assert true <= 1e-12


def test_gmres_exact_in_n_steps():
    """Full GMRES on an n x n system terminates with the exact solution in <= n steps."""
    rng = np.random.default_rng(4)
    n = 12
    A = rng.standard_normal((n, n)) + 3 * np.eye(n)
    b = rng.standard_normal(n)
    At = torch.from_numpy(A)
    x, its, _ = gmres(lambda v: At @ v, torch.from_numpy(b), tol=1e-14, restart=n, maxiter=1)
    assert its <= n
    assert np.abs(x.numpy() - np.linalg.solve(A, b)).max() < 1e-10


def test_newton_krylov_known_root():
    """F(z) = z^3 + A z - c with c built from a chosen root z*."""
    rng = np.random.default_rng(5)
    n = 30
    A = torch.from_numpy(np.eye(n) * 2 + 0.1 * rng.standard_normal((n, n)))
    zs = torch.from_numpy(rng.uniform(-1, 1, n))
    c = zs ** 3 + A @ zs

    def F(z):
        return z ** 3 + A @ z - c

    sol = newton_krylov(F, torch.zeros(n, dtype=torch.float64), fatol=1e-13, maxiter=50)
    assert sol.success, sol.message
    assert float((sol.x - zs).abs().max()) < 1e-12
    assert sol.fnorm <= 1e-13
    # quadratic convergence at the end
    h = sol.history
    assert h[-1] < 1e-13 and h[-2] < 1e-4


def test_newton_krylov_reports_failure():
    """No real root: the line search gives up and the result says so."""
    def F(z):
        return z * z + 1.0

    sol = newton_krylov(F, torch.ones(3, dtype=torch.float64), fatol=1e-10, maxiter=30)
    assert not sol.success
    assert sol.status in (1, 2)


# ---- the reference's steady state pins the oracle ---------------------------
def test_golden_state_is_a_fixed_point_of_the_oracle(steady):
    ns = oracle_ns()
    z = np.concatenate([steady["steady_Xs"], steady["steady_pres_s"].ravel()])
    load(ns, z)
    ns.update()
    F = (vec(ns, pres=True) - z) / ns.dt
    assert np.abs(F).max() < 1e-11
    assert steady["steady_res"][-1] < 1e-12


def test_oracle_steady_fun_matches_reference(steady):
    ns = oracle_ns()
    X0 = steady["steady_X0"]
    load(ns, X0)
    ns.pres.vhat = steady["steady_pres0"].copy()
    F = ns.steady_residual()
    scale = np.abs(X0).max() / ns.dt
    assert np.abs(F - steady["steady_F0"]).max() < 1e-13 * scale
    load(ns, X0)
    ns.pres.vhat = steady["steady_pres0"].copy()
    for _ in range(20):
        ns.update()
    F = (vec(ns) - X0) / 0.2
    assert np.abs(F - steady["steady_F0_dt"]).max() < 1e-12 * np.abs(X0).max() / 0.2


def test_newton_krylov_on_oracle_map_finds_reference_attractor(steady):
    """z = [T, U, V, pres], F = (Phi^50(z) - z)/0.5 on the oracle's step."""
    ns = oracle_ns()
    nst = 50

    def F(z):
        load(ns, z.numpy())
        for _ in range(nst):
            ns.update()
        return (torch.from_numpy(vec(ns, pres=True)) - z) / (nst * ns.dt)

    z0 = torch.from_numpy(np.concatenate([steady["steady_X0"], steady["steady_pres0"].ravel()]))
    sol = newton_krylov(F, z0, fatol=1e-11, maxiter=30)
    assert sol.success, sol.message
    nx = steady["steady_Xs"].size
    x = sol.x.numpy()
    Xs = steady["steady_Xs"]
    assert np.abs(x[:nx] - Xs).max() < 1e-9 * np.abs(Xs).max()
    assert np.abs(x[nx:] - steady["steady_pres_s"].ravel()).max() < 1e-7
    load(ns, x)
    nu = o.eval_Nu(ns.T, ns.deriv_field, Tbc=ns.Tbc, Lz=ns.y[-1] - ns.y[0])
    assert abs(nu - steady["steady_Nu"][0]) < 1e-9
