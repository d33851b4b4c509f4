"""The reference's diffusion drivers (diffusion/diff_1d.py, diffusion/diff_2d-bc.py)
written against the drop-in API, and the ChebDirichletNeumann basis, on the GPU.

The drivers are user scripts in the reference (classes defined in the script,
driven through Base, Field, FieldBC, grad, the Helmholtz plan templates and
Integrator); they are restated here as the smoke tests of those APIs:
  1-D: du/dt = kappa u'' + cos(pi x / 2), u(+-1) = 0, implicit Euler.  Its
       steady state is 4 cos(pi x / 2) / (pi^2 kappa) exactly; the discrete
       fixed point reaches it to spectral accuracy.
  2-D: du/dt = kappa lap u on ChebDirichlet x ChebNeumann with u(-1, y) =
       cos(pi y), u(1, y) = 0 lifted by FieldBC, Crank-Nicolson-type ADI
       (beta = 0.5).  Compared step by step with the oracle, and at late
       times with the Laplace solution cos(pi y) sinh(pi (1 - x)) / sinh(2 pi):
       the ADI fixed point differs from it by the splitting term, first order
       in dt.
"""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


def host(t):
    return t.detach().cpu().numpy()


def test_dirichlet_neumann_basis_matches_oracle():
    from oracle import pypde_oracle as o
    import pypde_amd as pd
    rng = np.random.default_rng(21)
    for N in (12, 33):
        b, ob = pd.Base(N, "CDN"), o.Basis(N, "CDN")
        assert b.id == "CDN" and b.M == N - 2
        assert np.array_equal(b.S, ob.S)
        v = rng.standard_normal((N - 2, 5))
        u = rng.standard_normal((N, 5))
        assert rel(host(b.to_chebyshev(v)), ob.to_chebyshev(v)) < 1e-15
        assert rel(host(b.from_chebyshev(u)), ob.from_chebyshev(u)) < 1e-13
        f = rng.standard_normal((N, 5))
        assert rel(host(b.forward_fft(f)), ob.forward_fft(f)) < 1e-13
        assert rel(host(b.backward_fft(v)), ob.backward_fft(v)) < 1e-13
        for d in (1, 2):
            assert rel(host(b.derivative(v, d)), ob.derivative(v, d)) < 1e-13
        # along axis 1 of a 2-D field
        F, oF = pd.Field([pd.Base(N, "CD"), pd.Base(N, "CDN")]), \
            o.Field([o.Basis(N, "CD"), o.Basis(N, "CDN")])
        g = rng.standard_normal((N, N))
        assert rel(host(F.forward(g)), oF.forward(g)) < 1e-13


class _Diffusion1d:
    """diffusion/diff_1d.py:9-96 on the drop-in API."""

    def __init__(self, N=50, kappa=1.0, dt=0.2, base="CD"):
        import pypde_amd as pd
        self.kappa, self.dt = kappa, dt
        self.field = pd.Field([pd.Base(N, base)])
        self.solver = pd.solverplan_hholtz1d(self.field.xs, lam=dt * kappa)
        f = np.cos(np.pi / 2 * self.field.x)
        self.fhat = self.field.xs[0].family.forward_fft(f)

    def update(self):
        rhs = self.solver.solve_rhs(self.dt * self.fhat)
        rhs = rhs + self.solver.solve_old(self.field.vhat)
        self.field.vhat = self.solver.solve_lhs(rhs)


def test_diffusion_1d_driver():
    D = _Diffusion1d(N=50, kappa=1.0, dt=0.5)
    # first steps against the dense restatement of the same plan
    from oracle import pypde_oracle as o
    b = o.Basis(50, "CD")
    S, B, I = b.S, b.family.B(2, 2), b.family.I(2)
    A = B @ S - D.dt * D.kappa * (I @ S)
    fhat = b.family.forward_fft(np.cos(np.pi / 2 * b.x))
    v = np.zeros(48)
    for _ in range(3):
        D.update()
        v = np.linalg.solve(A, B @ (D.dt * fhat) + B @ S @ v)
    assert rel(host(D.field.vhat), v) < 1e-12
    for _ in range(200):
        D.update()
    u = host(D.field.backward(D.field.vhat))
    exact = 4.0 * np.cos(np.pi * D.field.x / 2) / (np.pi ** 2 * D.kappa)
    assert np.abs(u - exact).max() < 1e-12


def _diffusion_2d(pd, N, dt, kappa, beta, nsteps, oracle=False):
    """diffusion/diff_2d-bc.py:9-107 (CD x CN, x-boundary values cos(pi y) at
    x = -1), on the drop-in API or on the oracle."""
    if oracle:
        Base = lambda n, k: pd.Basis(n, k)
        grad = lambda f, d: pd.grad(f, d)
        as_arr = lambda a: a
    else:
        Base = lambda n, k: pd.Base(n, k)
        grad = lambda f, d: pd.grad(f, deriv=d, return_field=True).vhat
        as_arr = host
    field = pd.Field([Base(N, "CD"), Base(N, "CN")])
    bc = np.zeros((2, N))
    bc[0, :] = np.cos(np.pi * field.y)
    fieldbc = pd.FieldBC(field.xs, axis=0)
    fieldbc.add_bc(bc)
    solver = pd.solverplan_hholtz2d_adi(field.xs, lam=dt * kappa * beta)
    fhat = dt * kappa * grad(fieldbc, (0, 2))
    history = []
    for _ in range(nsteps):
        rhs = fhat + dt * kappa * (1.0 - beta) * grad(field, (0, 2))
        rhs = rhs + dt * kappa * (1.0 - beta) * grad(field, (2, 0))
        r = solver.solve_rhs(rhs)
        r = r + solver.solve_old(field.vhat)
        field.vhat = solver.solve_lhs(r)
        history.append(np.array(as_arr(field.vhat)))
    return field, fieldbc, history


def test_diffusion_2d_bc_driver():
    import pypde_amd as pd
    from oracle import pypde_oracle as o
    N, dt, kappa, beta = 24, 0.01, 0.5, 0.5
    field, fieldbc, hist = _diffusion_2d(pd, N, dt, kappa, beta, 40)
    _, _, ohist = _diffusion_2d(o, N, dt, kappa, beta, 40, oracle=True)
    for k in (0, 9, 39):
        assert rel(hist[k], ohist[k]) < 1e-12, k
    # late time: the ADI fixed point approaches the Laplace solution at first
    # order in dt (the splitting term); the oracle gives 1.10e-2 at dt = 0.05
    # and 2.25e-3 at dt = 0.01
    errs = []
    for dt_, n in ((0.05, 600), (0.01, 3000)):
        field, fieldbc, hist = _diffusion_2d(pd, N, dt_, kappa, beta, n)
        total = host(field.backward(field.vhat)) + host(fieldbc.backward(fieldbc.vhat))
        x, y = np.meshgrid(field.x, field.y, indexing="ij")
        exact = np.cos(np.pi * y) * np.sinh(np.pi * (1 - x)) / np.sinh(2 * np.pi)
        errs.append(np.abs(total - exact).max())
        assert np.abs(hist[-1] - hist[-2]).max() < 1e-14  # steady
    assert errs[0] < 1.2e-2 and 4.5 < errs[0] / errs[1] < 5.5, errs
