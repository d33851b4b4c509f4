"""Write the golden fixtures for the functions that no closed form pins.

Run where the reference source tree is installed (machines that only run the
tests do not need it):  python tests/golden/make_golden.py

Imports preiter93/pypde from /root/reference to record input/output vectors of
  * navier.rbc2d.NavierStokes.update   (navier/rbc2d.py:396-434): single-step
    states and Nusselt / kinetic-energy trajectories,
  * navier.rbc2d_adj.NavierStokesAdjoint.update (navier/rbc2d_adj.py:321-356),
  * NavierStokesSteadyState.steady_fun (navier/rbc2d_base.py:321-341).
Everything else on the path (transforms, stencils, banded solves, Helmholtz,
Poisson) is pinned by closed forms in tests/test_oracle_*.py, so the reference
is not run for those.

How the import is made to work without loading anything prebuilt:
  * the four f2py extension modules are served by oracle/_ref/libpypde_f90.so,
    i.e. the reference's own .f90 sources compiled here with flang
    (oracle/Makefile), behind adapters that keep f2py's calling convention
    (intent(out) zero-filled, intent(inout) in place on Fortran-contiguous
    arrays).  The shipped .so files are never opened: the module names are
    placed in sys.modules before the reference package is imported;
  * h5py (file IO only, not on the path) is an empty module;
  * pypde.plot.initplot (matplotlib styling run at import) is a no-op.
"""
import ctypes
import os
import sys
import time
import types

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(REPO, "oracle", "_ref", "libpypde_f90.so")

_dp = ctypes.POINTER(ctypes.c_double)
_i = ctypes.c_int


def _load_f90():
    if not os.path.exists(LIB):
        raise SystemExit("build oracle/_ref first: make -C oracle ref")
    lib = ctypes.CDLL(LIB)
    lib.ref_diff_2d.argtypes = [_dp, _dp, _i, _i]
    lib.ref_tdma_2d.argtypes = [_dp, _dp, _dp, _dp, _i, _dp, _i, _i]
    lib.ref_fdma_2d.argtypes = [_dp, _dp, _dp, _dp, _dp, _i, _i, _i]
    lib.ref_init_fdma.argtypes = [_dp, _dp, _dp, _dp, _dp, _i]
    lib.ref_fdma_type2.argtypes = [_dp, _dp, _dp, _dp, _i, _i, _i, _i]
    lib.ref_twodma_2d.argtypes = [_dp, _dp, _dp, _i, _i, _i]
    return lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _fin(a):
    return np.asfortranarray(a, dtype=np.float64)


def _inout(x):
    if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.f_contiguous):
        raise ValueError("failed to initialize intent(inout) array -- input not fortran contiguous")
    return x


def _install_f2py_modules(lib):
    def diff_1d(c):
        c = _fin(c)
        dc = np.zeros((c.shape[0], 1), order="F")
        lib.ref_diff_2d(_p(_fin(c[:, None])), _p(dc), c.shape[0], 1)
        return dc[:, 0].copy()

    def diff_2d(c):
        c = _fin(c)
        dc = np.zeros(c.shape, order="F")
        lib.ref_diff_2d(_p(c), _p(dc), c.shape[0], c.shape[1])
        return dc

    def solve_tdma_1d(a, b, c, d, k):
        d2 = _fin(np.asarray(d)[:, None])
        x = np.zeros(d2.shape, order="F")
        lib.ref_tdma_2d(_p(_fin(a)), _p(_fin(b)), _p(_fin(c)), _p(d2), int(k), _p(x),
                        d2.shape[0], 1)
        return x[:, 0].copy()

    def solve_tdma_2d(a, b, c, d, k):
        d2 = _fin(d)
        x = np.zeros(d2.shape, order="F")
        lib.ref_tdma_2d(_p(_fin(a)), _p(_fin(b)), _p(_fin(c)), _p(d2), int(k), _p(x),
                        d2.shape[0], d2.shape[1])
        return x

    def solve_fdma_1d(l, d, u1, u2, x):
        x = _inout(x)
        lib.ref_fdma_2d(_p(_fin(l)), _p(_fin(d)), _p(_fin(u1)), _p(_fin(u2)), _p(x), 0,
                        x.shape[0], 1)

    def solve_fdma_2d(l, d, u1, u2, x, axis):
        x = _inout(x)
        lib.ref_fdma_2d(_p(_fin(l)), _p(_fin(d)), _p(_fin(u1)), _p(_fin(u2)), _p(x),
                        int(axis), x.shape[0], x.shape[1])

    def init_fdma(A):
        A = _fin(A)
        n = A.shape[0]
        d, u1, u2, l = np.zeros(n), np.zeros(n - 2), np.zeros(max(n - 4, 1)), np.zeros(n - 2)
        lib.ref_init_fdma(_p(A), _p(d), _p(u1), _p(u2), _p(l), n)
        return d, u1, u2[: n - 4], l

    def solve_fdma_type2(A, C, lam, x, axis, singular):
        x = _inout(x)
        lib.ref_fdma_type2(_p(_fin(A)), _p(_fin(C)), _p(_fin(lam)), _p(x), int(axis),
                           int(bool(singular)), x.shape[0], x.shape[1])

    def solve_twodma_1d(d, u, x):
        x = _inout(x)
        lib.ref_twodma_2d(_p(_fin(d)), _p(_fin(u)), _p(x), 0, x.shape[0], 1)

    def solve_twodma_2d(d, u1, x, axis):
        x = _inout(x)
        lib.ref_twodma_2d(_p(_fin(d)), _p(_fin(u1)), _p(x), int(axis), x.shape[0], x.shape[1])

    def mod(name, **fns):
        m = types.ModuleType(name)
        m.__dict__.update(fns)
        sys.modules[name] = m
        return m

    pk = [mod("pypde.bases.fortran"), mod("pypde.bases.linalg.fortran"),
          mod("pypde.solver.linalg.fortran")]
    for p in pk:
        p.__path__ = []
    pk[0].differentiate_cheby = mod("pypde.bases.fortran.differentiate_cheby",
                                    diff_1d=diff_1d, diff_2d=diff_2d)
    pk[1].tdma = mod("pypde.bases.linalg.fortran.tdma", solve_tdma_1d=solve_tdma_1d,
                     solve_tdma_2d=solve_tdma_2d)
    pk[2].fdma = mod("pypde.solver.linalg.fortran.fdma", solve_fdma_1d=solve_fdma_1d,
                     solve_fdma_2d=solve_fdma_2d, init_fdma=init_fdma,
                     solve_fdma_type2=solve_fdma_type2)
    pk[2].twodma = mod("pypde.solver.linalg.fortran.twodma", solve_twodma_1d=solve_twodma_1d,
                       solve_twodma_2d=solve_twodma_2d)


def import_reference():
    lib = _load_f90()
    sys.modules.setdefault("h5py", types.ModuleType("h5py"))
    _install_f2py_modules(lib)
    sys.path.insert(0, REF)
    import matplotlib
    matplotlib.use("Agg")
    import pypde
    import pypde.plot as _plot

    def _noop(*a, **k):
        return None

    pypde.initplot = _noop
    _plot.initplot = _noop
    sys.modules["pypde.plot.initplot"].initplot = _noop
    from navier.rbc2d import NavierStokes
    from navier.rbc2d_adj import NavierStokesAdjoint
    from pypde.field_operations import avg_vol
    return NavierStokes, NavierStokesAdjoint, avg_vol


def seeded(ns, seed, amplitude=1e-3):
    """Same convention as oracle.pypde_oracle.seeded_initial_state."""
    rng = np.random.default_rng(seed)
    ns.T.v = rng.uniform(-amplitude, amplitude, size=ns.T.v.shape)
    ns.T.forward()


def kinetic_energy(ns, avg_vol):
    u = ns.U.backward(ns.U.vhat)
    v = ns.V.backward(ns.V.vhat)
    f = ns.deriv_field
    return 0.5 * avg_vol(u * u + v * v, f.dx, f.dy)


def quiet_nu(ns):
    import io
    import contextlib
    with contextlib.redirect_stdout(io.StringIO()):
        return ns.eval_Nu()


def state(ns):
    return dict(T=ns.T.vhat.copy(), U=ns.U.vhat.copy(), V=ns.V.vhat.copy(),
                P=ns.P.vhat.copy(), pres=ns.pres.vhat.copy())


def main():
    NavierStokes, NavierStokesAdjoint, avg_vol = import_reference()
    out = {}

    # 1) single steps from a seeded state with nonzero velocity (small, non-square)
    cases = [
        ("linear_eu", dict(case="linear", integrator="eu", beta=1.0)),
        ("rbc_eu", dict(case="rbc", integrator="eu", beta=1.0)),
        ("rbc_rk3", dict(case="rbc", integrator="rk3", beta=1.0)),
        ("rbc_eu_beta", dict(case="rbc", integrator="eu", beta=0.5)),
        ("zero_eu", dict(case="zero", integrator="eu", beta=1.0)),
    ]
    for name, kw in cases:
        case = kw.pop("case")
        ns = NavierStokes(case=case, shape=(20, 16), ra=1e5, pr=1.0, dt=0.01, tsave=None,
                          dealias=True, aspect=1.0, **kw)
        seeded(ns, seed=3, amplitude=0.1)
        rng = np.random.default_rng(4)
        ns.U.vhat[:] = 1e-2 * rng.standard_normal(ns.U.vhat.shape)
        ns.V.vhat[:] = 1e-2 * rng.standard_normal(ns.V.vhat.shape)
        ns.pres.vhat[:] = 1e-3 * rng.standard_normal(ns.pres.vhat.shape)
        s0 = state(ns)
        for k, v in s0.items():
            out[f"step_{name}_in_{k}"] = v
        for n in range(3):
            ns.update()
            ns.update_time()
            if n == 0:
                for k, v in state(ns).items():
                    out[f"step_{name}_out1_{k}"] = v
        for k, v in state(ns).items():
            out[f"step_{name}_out3_{k}"] = v

    # 2) configuration c1: 64x64, Ra=1e5, Pr=1, dt=0.01, Euler, 1000 steps, seed 0
    for case in ("linear", "rbc"):
        ns = NavierStokes(case=case, shape=(64, 64), ra=1e5, pr=1.0, dt=0.01, tsave=None,
                          dealias=True, integrator="eu", beta=1.0, aspect=1.0)
        seeded(ns, seed=0)
        nsteps = 1000
        nu_plate, nu_vol, ke = [], [], []
        t0 = time.perf_counter()
        for n in range(nsteps):
            ns.update()
            ns.update_time()
            a, b = quiet_nu(ns)
            nu_plate.append(a)
            nu_vol.append(b)
            ke.append(kinetic_energy(ns, avg_vol))
        dt_wall = time.perf_counter() - t0
        print(f"c1 {case}: {nsteps} steps in {dt_wall:.1f} s")
        out[f"c1_{case}_nu_plate"] = np.array(nu_plate)
        out[f"c1_{case}_nu_vol"] = np.array(nu_vol)
        out[f"c1_{case}_ke"] = np.array(ke)
        for k, v in state(ns).items():
            out[f"c1_{case}_final_{k}"] = v

    # 3) steady residual (rbc2d_base.py:321-341) at a small size
    ns = NavierStokes(case="rbc", shape=(20, 16), ra=5e4, pr=1.0, dt=0.005, tsave=None)
    seeded(ns, seed=7, amplitude=0.1)
    X = ns.vectorify()
    out["steady_X"] = X
    out["steady_res"] = ns.steady_fun(X.copy(), ns, None)

    # 4) adjoint-descent step (rbc2d_adj.py:321-356) at a small size
    nsa = NavierStokesAdjoint(case="rbc", shape=(20, 16), ra=5e4, pr=1.0, dt=0.005,
                              tsave=None, dealias=True, integrator="eu", beta=1.0)
    rng = np.random.default_rng(9)
    nsa.T.vhat[:] = 1e-1 * rng.standard_normal(nsa.T.vhat.shape)
    nsa.U.vhat[:] = 1e-2 * rng.standard_normal(nsa.U.vhat.shape)
    nsa.V.vhat[:] = 1e-2 * rng.standard_normal(nsa.V.vhat.shape)
    out["adj_in_T"], out["adj_in_U"], out["adj_in_V"] = (
        nsa.T.vhat.copy(), nsa.U.vhat.copy(), nsa.V.vhat.copy())
    nsa.update()
    out["adj_out_T"], out["adj_out_U"], out["adj_out_V"] = (
        nsa.T.vhat.copy(), nsa.U.vhat.copy(), nsa.V.vhat.copy())
    out["adj_out_pres"] = nsa.pres.vhat.copy()
    out["adj_ns_pres"] = nsa.NS.pres.vhat.copy()

    path = os.path.join(HERE, "rbc2d_reference.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


def c3(nsteps=2000):
    """Configuration c3: 512x512, rbc, Ra=1e8, Pr=1, dt=1e-3, Euler, seed 0, noise
    1e-3, 2000 steps: Nusselt / kinetic-energy series, the leading 32x32 block
    of the final temperature coefficients and the reference's wall time per step
    (one thread)."""
    NavierStokes, _, avg_vol = import_reference()
    ns = NavierStokes(case="rbc", shape=(512, 512), ra=1e8, pr=1.0, dt=1e-3, tsave=None,
                      dealias=True, integrator="eu", beta=1.0, aspect=1.0)
    seeded(ns, seed=0)
    nu_plate, nu_vol, ke, wall = [], [], [], []
    for n in range(nsteps):
        t0 = time.perf_counter()
        ns.update()
        ns.update_time()
        wall.append(time.perf_counter() - t0)
        a, b = quiet_nu(ns)
        nu_plate.append(a)
        nu_vol.append(b)
        ke.append(kinetic_energy(ns, avg_vol))
        if n % 50 == 0:
            print("c3 step %d  %.3f s/step  Nu %.12f  KE %.6e" % (n, wall[-1], a, ke[-1]), flush=True)
    out = dict(c3_nu_plate=np.array(nu_plate), c3_nu_vol=np.array(nu_vol), c3_ke=np.array(ke),
               c3_T_final_32=ns.T.vhat[:32, :32].copy(), c3_sec_per_step=np.array(wall))
    path = os.path.join(HERE, "rbc2d_c3_512.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


def steady():
    """Steady state of the rbc time-step map, 20x16, Ra=1e4, Pr=1, dt=0.01.

    The reference's solve_steady_state (navier/rbc2d_base.py:259-291) does not
    converge on this case (its steady_fun, navier/rbc2d_base.py:318-338, carries
    the pressure from call to call, so the residual is not a function of X), so
    the fixed point is taken from the map itself: the reference is stepped from
    set_temperature(0.2) until max|F| stops falling, F = (step(X) - X)/dt.  Also
    stores steady_fun at the state after 300 steps, with dt=None (one step) and
    dt=0.2 (iterate over 0.2 time units), each from the same pressure."""
    NavierStokes, _, _ = import_reference()
    ns = NavierStokes(case="rbc", shape=(20, 16), ra=1e4, pr=1.0, dt=0.01, tsave=None,
                      dealias=True, integrator="eu", beta=1.0, aspect=1.0)
    ns.set_temperature(amplitude=0.2)
    for _ in range(300):
        ns.update()
    X0 = ns.vectorify()
    pres0 = ns.pres.vhat.copy()
    out = dict(steady_X0=X0, steady_pres0=pres0)
    out["steady_F0"] = ns.steady_fun(X0.copy(), ns, None)
    ns.pres.vhat[:] = pres0
    out["steady_F0_dt"] = ns.steady_fun(X0.copy(), ns, 0.2)
    ns.pres.vhat[:] = pres0
    T, U, V = ns.reshape(X0)
    ns.T.vhat[:], ns.U.vhat[:], ns.V.vhat[:] = T, U, V
    best, steps, hist = np.inf, 300, []
    while steps < 200000:
        for _ in range(999):
            ns.update()
        X = ns.vectorify()
        ns.update()
        steps += 1000
        r = np.abs(ns.vectorify() - X).max() / ns.dt
        hist.append(r)
        if r > 0.7 * best and r < 1e-9:
            break
        best = min(best, r)
    out["steady_Xs"] = ns.vectorify()
    out["steady_pres_s"] = ns.pres.vhat.copy()
    out["steady_res"] = np.array(hist)
    out["steady_steps"] = np.array(steps)
    out["steady_Nu"] = np.array(quiet_nu(ns))
    print("steady: steps", steps, "res", hist[-3:], "Nu", out["steady_Nu"])
    path = os.path.join(HERE, "rbc2d_steady.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "c3":
        c3()
    elif len(sys.argv) > 1 and sys.argv[1] == "steady":
        steady()
    else:
        main()
