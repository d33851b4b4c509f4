"""The step-level C-ABI: pypde_rbc_step runs a recorded step plan on buffers
the caller owns, through ctypes only (navier/rbc2d.py:396-434 for the step it
replaces), against the reference's own single steps (tests/golden)."""
import ctypes

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

CASES = {
    "linear_eu": dict(case="linear", integrator="eu", beta=1.0),
    "rbc_eu": dict(case="rbc", integrator="eu", beta=1.0),
    "zero_eu": dict(case="zero", integrator="eu", beta=1.0),
    "rbc_rk3": dict(case="rbc", integrator="rk3", beta=1.0),
    "rbc_eu_beta": dict(case="rbc", integrator="eu", beta=0.5),
}


def host(t):
    return t.detach().cpu().numpy()


def _engine(name):
    from pypde_amd.navier import NavierStokes
    kw = dict(CASES[name])
    case = kw.pop("case")
    ns = NavierStokes(case=case, shape=(20, 16), ra=1e5, pr=1.0, dt=0.01, tsave=None,
                      dealias=True, aspect=1.0, **kw)
    assert ns.fused
    return ns.engine


@pytest.mark.parametrize("name", sorted(CASES))
def test_rbc_step_on_caller_buffers_matches_reference(golden, name):
    """The plan recorded on the engine's buffers, run by pypde_rbc_step on other
    state arrays and a workspace filled with NaN (every workspace entry the
    step reads is written first): one step and then two more against the
    reference's states, and bit for bit the engine's own step."""
    import torch
    from pypde_amd._lib import lib
    eng = _engine(name)
    plan = eng.build_plan()
    h = plan.handle
    wsb = lib.pypde_plan_region_bytes(h, 3)
    assert 0 < wsb <= 8 * eng.ws.numel()
    assert lib.pypde_plan_region_bytes(h, 0) == 8 * eng.X.numel()
    pre = "step_%s_" % name
    dev = eng.X.device
    X = torch.empty_like(eng.X)
    pres = torch.empty_like(eng.pres)
    P = torch.zeros_like(eng.P)
    for i, f in enumerate("UVT"):
        X[0, i] = torch.as_tensor(golden[pre + "in_" + f], device=dev)
    pres[0] = torch.as_tensor(golden[pre + "in_pres"], device=dev)
    ws = torch.full((wsb // 8,), float("nan"), dtype=torch.float64, device=dev)
    # the engine's own buffers start from the same state
    eng.X.copy_(X)
    eng.pres.copy_(pres)
    eng.P.zero_()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    assert lib.pypde_rbc_step(h, vp(X), vp(pres), vp(P), vp(ws), 1, s) == 0
    torch.cuda.synchronize()
    for i, f in enumerate("UVT"):
        e = rel(host(X[0, i]), golden[pre + "out1_" + f])
        assert e < 1e-12, (name, f, e)
    assert rel(host(pres[0]), golden[pre + "out1_pres"]) < 1e-12
    assert lib.pypde_rbc_step(h, vp(X), vp(pres), vp(P), vp(ws), 2, s) == 0
    for _ in range(3):
        eng.step_eager()
    torch.cuda.synchronize()
    for i, f in enumerate("UVT"):
        e = rel(host(X[0, i]), golden[pre + "out3_" + f])
        assert e < 3e-12, (name, f, e)
    assert np.array_equal(host(X), host(eng.X))
    assert np.array_equal(host(pres), host(eng.pres))
    assert np.array_equal(host(P), host(eng.P))


def test_rbc_step_errors_are_reported():
    import torch
    from pypde_amd._lib import lib
    eng = _engine("rbc_eu")
    h = eng.build_plan().handle
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    assert lib.pypde_rbc_step(h, vp(eng.X), None, vp(eng.P), vp(eng.ws), 1, s) != 0
    assert b"null" in lib.pypde_last_error()
    assert lib.pypde_rbc_step(h, vp(eng.X), vp(eng.pres), vp(eng.P), vp(eng.ws), -1, s) != 0
    assert b"negative" in lib.pypde_last_error()


def test_batched_plan_steps_every_instance():
    """A plan of several instances (per-instance Ra, dt) run through the C-ABI
    equals the same instances stepped one per plan (1e-11: the batched and
    single-instance plans may use different GEMM kernels, whose rounding the
    Helmholtz solves amplify)."""
    import torch
    from pypde_amd._lib import lib
    from pypde_amd.engine import RBCBatch
    params = [dict(ra=1e4, pr=1.0, dt=0.01), dict(ra=3e5, pr=1.0, dt=0.004),
              dict(ra=2e6, pr=0.7, dt=0.002)]
    shape = (24, 20)
    eng = RBCBatch(shape, params, case="rbc")
    singles = [RBCBatch(shape, [p], case="rbc") for p in params]
    rng = np.random.default_rng(4)
    X0 = torch.as_tensor(rng.standard_normal(tuple(eng.X.shape)) * 0.05, device=eng.X.device)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    X, pres, P = X0.clone(), torch.zeros_like(eng.pres), torch.zeros_like(eng.P)
    ws = torch.empty(lib.pypde_plan_region_bytes(eng.build_plan().handle, 3) // 8,
                     dtype=torch.float64, device=X.device)
    assert lib.pypde_rbc_step(eng.plan.handle, vp(X), vp(pres), vp(P), vp(ws), 4, s) == 0
    for b, e1 in enumerate(singles):
        e1.X[0].copy_(X0[b])
        for _ in range(4):
            e1.step_eager()
        torch.cuda.synchronize()
        assert rel(host(X[b]), host(e1.X[0])) < 1e-11, b
        assert rel(host(pres[b]), host(e1.pres[0])) < 1e-11, b


@pytest.mark.parametrize("name", ["rbc_eu", "rbc_rk3", "rbc_eu_beta"])
def test_step_vjp_plan_on_caller_buffers(name):
    """pypde_rbc_step_vjp (the step's VJP recorded as a plan; three-stage
    states recomputed inside its workspace) on other input and cotangent
    arrays and a NaN-filled workspace gives bit for bit StepAdjoint.vjp, whose
    launches it replays; the engine's own state is not touched."""
    import torch
    from pypde_amd._lib import lib
    from pypde_amd.adjoint import StepAdjoint
    eng = _engine(name)
    adj = StepAdjoint(eng)
    rng = np.random.default_rng(11)
    dev = eng.X.device
    rnd = lambda t, a: torch.as_tensor(rng.standard_normal(tuple(t.shape)) * a, device=dev)
    X, pres = rnd(eng.X, 0.05), rnd(eng.pres, 0.01)
    Xb, pb = rnd(eng.X, 1.0), rnd(eng.pres, 1.0)
    ref_X, ref_p = adj.vjp(X, pres, Xb, pb)
    eng_state = [t.clone() for t in (eng.X, eng.pres)]
    plan = adj.build_plan()
    wsb = lib.pypde_rbc_step_vjp_workspace_bytes(plan.handle)
    assert wsb > 0
    ws = torch.full((wsb // 8 + 32,), float("nan"), dtype=torch.float64, device=dev)
    off = (-ws.data_ptr()) % 256 // 8
    wsa = ws[off:]
    Xc, pc, Xbc, pbc = X.clone(), pres.clone(), Xb.clone(), pb.clone()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    assert lib.pypde_rbc_step_vjp(plan.handle, vp(Xc), vp(pc), vp(Xbc), vp(pbc), vp(wsa), s) == 0
    torch.cuda.synchronize()
    assert np.array_equal(host(Xbc), host(ref_X)), name
    assert np.array_equal(host(pbc), host(ref_p)), name
    assert np.array_equal(host(Xc), host(X)) and np.array_equal(host(pc), host(pres))
    for t, v in zip((eng.X, eng.pres), eng_state):
        assert np.array_equal(host(t), host(v))
    assert np.isfinite(host(Xbc)).all()
    assert lib.pypde_rbc_step_vjp(plan.handle, vp(Xc), None, vp(Xbc), vp(pbc), vp(wsa), s) != 0
    assert b"null" in lib.pypde_last_error()
