"""Pin the oracle's bases, transforms and plans against closed forms (CPU)."""
import numpy as np
import numpy.polynomial.chebyshev as npcheb
import pytest

from oracle import pypde_oracle as o
from conftest import rel

rng = np.random.default_rng(5)


def test_grid_closed_form():
    for N in (5, 16, 33):
        x = o.gauss_lobatto(N - 1)
        assert np.allclose(x, -np.cos(np.pi * np.arange(N) / (N - 1)), atol=1e-15)
        assert np.all(np.diff(x) > 0)


@pytest.mark.parametrize("N", [8, 24, 41])
def test_chebyshev_backward_is_series_evaluation(N):
    b = o.Basis(N, "CH")
    c = rng.standard_normal((N, 3))
    u = b.backward_fft(c)
    ref = np.stack([npcheb.chebval(b.x, c[:, j]) for j in range(3)], axis=1)
    assert rel(u, ref) < 1e-13
    # forward is the exact inverse at the Gauss-Lobatto nodes
    assert rel(b.forward_fft(u), c) < 1e-13


def test_forward_brute_force_sum():
    N = 12
    b = o.Basis(N, "CH")
    f = rng.standard_normal(N)
    c = b.forward_fft(f)
    k = np.arange(N)
    ref = np.zeros(N)
    for kk in range(N):
        s = f[0] + (-1) ** kk * f[-1] + 2 * np.sum(f[1:-1] * np.cos(np.pi * np.arange(1, N - 1) * kk / (N - 1)))
        w = 1.0 if kk in (0, N - 1) else 2.0
        ref[kk] = w * (-1) ** kk * s / (2 * (N - 1))
    assert np.allclose(c, ref, atol=1e-14)


@pytest.mark.parametrize("key", ["CD", "CN"])
def test_galerkin_maps(key):
    N = 30
    b = o.Basis(N, key)
    S = b.S
    v = rng.standard_normal(N - 2)
    u = b.to_chebyshev(v)
    assert np.allclose(u, S @ v, atol=1e-15)
    # the least-squares map of chebyshev.py:295-337 equals lstsq on S
    uu = rng.standard_normal(N)
    vv = b.from_chebyshev(uu)
    assert rel(vv, np.linalg.lstsq(S, uu, rcond=None)[0]) < 1e-12
    # boundary conditions of the basis functions
    phi = S.T  # rows: chebyshev coefficients of phi_k
    if key == "CD":
        assert np.allclose(phi @ np.ones(N), 0) and np.allclose(phi @ (-1.0) ** np.arange(N), 0)
    else:
        k2 = np.arange(N) ** 2.0
        assert np.allclose(phi @ k2, 0, atol=1e-12)


def test_pseudoinverse_identity():
    N = 20
    B = o.pseudoinverse_d2(N)
    D2 = o.diff_mat_spectral(N, 2)
    # B D2 = I on the kept rows 2: (the pseudo-inverse property); a wrong sign or
    # shifted diagonal in B breaks it
    assert np.allclose((B @ D2)[2:, 2:], np.eye(N - 2), atol=1e-12)


def _kron_helmholtz(field, lam, scale):
    bx, by = field.xs
    _, _, Ax = o.hholtz_matrices(bx, lam, scale[0])
    _, _, Ay = o.hholtz_matrices(by, lam, scale[1])
    return Ax, Ay


def test_helmholtz_adi_matches_kronecker():
    shape, lam, scale = (14, 12), 0.3, (0.5, 0.5)
    f = o.Field([o.Basis(shape[0], "CD"), o.Basis(shape[1], "CN")])
    plan = o.solverplan_hholtz2d_adi(f.xs, lam, scale)
    R = rng.standard_normal(shape)
    V0 = rng.standard_normal((shape[0] - 2, shape[1] - 2))
    rhs = plan.solve_rhs(R) + plan.solve_old(V0)
    X = plan.solve_lhs(rhs)
    Ax, Ay = _kron_helmholtz(f, lam, scale)
    m = plan.mats
    rhs_ref = m["Bx"] @ R @ m["By"].T + m["BSx"] @ V0 @ m["BSy"].T
    Xref = np.linalg.solve(np.kron(Ax, Ay), rhs_ref.ravel()).reshape(X.shape)
    assert rel(X, Xref) < 1e-12


def test_helmholtz_manufactured_solution():
    """test_hholtz2d.py:62-112: (1 - lam D2) u = f with u = cos(pi x/2) sin(pi y/2)."""
    shape, lam = (30, 30), 1 / np.pi ** 6
    f = o.Field([o.Basis(shape[0], "CD"), o.Basis(shape[1], "CN")])
    xx, yy = np.meshgrid(f.x, f.y, indexing="ij")
    sol = np.cos(np.pi / 2 * xx) * np.sin(np.pi / 2 * yy)
    rhs_f = (1.0 + 2 * lam * np.pi ** 2 / 4) * sol
    fc = o.Field([o.Basis(shape[0], "CH"), o.Basis(shape[1], "CH")])
    fhat = fc.forward(rhs_f)
    plan = o.solverplan_hholtz2d_adi(f.xs, lam)
    vhat = plan.solve_lhs(plan.solve_rhs(fhat))
    u = f.backward(vhat)
    assert np.abs(u - sol).max() < 1e-3  # ADI splitting error O(lam^2)


def test_poisson_matches_kronecker_with_gauge():
    shape, scale = (12, 10), (0.5, 0.5)
    f = o.Field([o.Basis(shape[0], "CN"), o.Basis(shape[1], "CN")])
    plan = o.solverplan_poisson2d(f.xs, singular=True, scale=scale)
    m = plan.mats
    R = rng.standard_normal(shape)
    X = plan.solve_lhs(plan.solve_rhs(R))
    X[0, 0] = 0
    # Ax P Cy^T + Cx P Ay^T = Bx R By^T on everything but the null mode
    L = np.kron(m["Ax"], m["Cy"]) + np.kron(m["Cx"], m["Ay"])
    lhs = (L @ X.ravel()).reshape(X.shape)
    rhs = m["Bx"] @ R @ m["By"].T
    # the constant mode (0,0) row is singular; compare the rest
    mask = np.ones_like(lhs, bool)
    mask[0, 0] = False
    assert rel(lhs[mask], rhs[mask]) < 1e-10


def test_poisson_dirichlet_manufactured():
    """test_poisson2d.py:81-120: D2 u = cos(a x) cos(a y), u = -cos cos / (2 a^2)."""
    shape, arg = (30, 24), np.pi / 2
    f = o.Field([o.Basis(shape[0], "CD"), o.Basis(shape[1], "CD")])
    plan = o.solverplan_poisson2d(f.xs, singular=False)
    xx, yy = np.meshgrid(f.x, f.y, indexing="ij")
    fc = o.Field([o.Basis(shape[0], "CH"), o.Basis(shape[1], "CH")])
    fhat = fc.forward(np.cos(arg * xx) * np.cos(arg * yy))
    u = f.backward(plan.solve_lhs(plan.solve_rhs(fhat)))
    assert np.abs(u - np.cos(arg * xx) * np.cos(arg * yy) * -1 / arg ** 2 / 2).max() < 1e-8


def test_dealias_product_matches_undealiased():
    """test_dealias.py:33-47."""
    N, M = 40, 20
    A = o.Field([o.Basis(N, "CD", 1.5), o.Basis(M, "CN", 1.5)])
    xx, yy = np.meshgrid(A.x, A.y, indexing="ij")
    a = np.sin(np.pi * xx) * np.cos(np.pi * yy)
    ah = A.forward(a)
    bh = A.forward(-a)
    ad = A.dealias.backward(ah)
    bd = A.dealias.backward(bh)
    assert ad.shape == (60, 30)
    chat = A.dealias.forward(ad * bd)
    assert chat.shape == ah.shape
    assert np.allclose(chat, A.forward(a * -a), rtol=1e-3, atol=1e-8)


@pytest.mark.parametrize("case", ["linear", "rbc"])
def test_conductive_state_has_unit_nusselt(case):
    # the purely conductive profile carries Nu = 1 exactly (plate and volume)
    ns = o.NavierStokes(case=case, shape=(16, 16), ra=1e4, pr=1.0, dt=0.01, tsave=None)
    nu0 = ns.eval_Nu()
    assert abs(nu0[0] - 1.0) < 1e-12 and abs(nu0[1] - 1.0) < 1e-12
    assert ns.eval_KE() == 0.0


def test_dirichlet_neumann_basis():
    """ChebDirichletNeumann (chebyshev.py:454-530): every basis function has
    phi(-1) = 0 and phi'(1) = 0; from_chebyshev inverts to_chebyshev; a
    function with those boundary values is expanded to spectral accuracy."""
    N = 30
    b = o.Basis(N, "CDN")
    S = b.S
    assert S.shape == (N, N - 2)
    for k in range(N - 2):
        assert abs(npcheb.chebval(-1.0, S[:, k])) < 1e-14
        assert abs(npcheb.chebval(1.0, npcheb.chebder(S[:, k]))) < 1e-12
    # the diagonals match the closed form T_k + d1 T_k+1 + d2 T_k+2
    k = 5.0
    den = 2 * k * k + 6 * k + 5
    assert S[5, 5] == 1.0 and abs(S[6, 5] - 4 * (k + 1) / den) < 1e-16
    assert abs(S[7, 5] + (2 * k * k + 2 * k + 1) / den) < 1e-16
    v = rng.standard_normal(N - 2)
    assert rel(b.from_chebyshev(b.to_chebyshev(v)), v) < 1e-13
    f = np.sin(np.pi * (b.x + 1) / 4)  # f(-1) = 0, f'(1) = 0
    vhat = b.forward_fft(f)
    assert np.abs(b.backward_fft(vhat) - f).max() < 1e-13
