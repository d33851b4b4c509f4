"""pypde_rbc_step_noalloc, the step entry point for callers that cannot
allocate (the plan's workspace is the library's device scratch array), against
pypde_rbc_step with a caller workspace: bit for bit, on two calls in a row, the
second on a smaller plan; and a plan whose workspace exceeds the scratch array
is rejected."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state(eng, seed):
    import torch
    rng = np.random.default_rng(seed)
    X = torch.as_tensor(rng.standard_normal(tuple(eng.X.shape)) * 1e-2, device=eng.X.device)
    pres = torch.as_tensor(rng.standard_normal(tuple(eng.pres.shape)) * 1e-3, device=eng.X.device)
    return X, pres, torch.zeros_like(eng.P)


def _compare(eng, seed, nsteps):
    import torch
    from pypde_amd._lib import lib
    h = eng.build_plan().handle
    wsb = lib.pypde_plan_region_bytes(h, 3)
    assert 0 < wsb <= lib.pypde_scratch_bytes()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    X, pres, P = _state(eng, seed)
    Xa, pa, Pa = X.clone(), pres.clone(), P.clone()
    ws = torch.full((wsb // 8,), float("nan"), dtype=torch.float64, device=X.device)
    assert lib.pypde_rbc_step(h, vp(Xa), vp(pa), vp(Pa), vp(ws), nsteps, s) == 0
    assert lib.pypde_rbc_step_noalloc(h, vp(X), vp(pres), vp(P), nsteps, s) == 0, \
        lib.pypde_last_error()
    torch.cuda.synchronize()
    assert torch.isfinite(X).all()
    assert torch.equal(X, Xa) and torch.equal(pres, pa) and torch.equal(P, Pa)


def test_noalloc_step_matches_step_with_workspace():
    from pypde_amd.engine import RBCBatch
    big = RBCBatch((256, 256), [dict(ra=1e8, pr=1.0, dt=1e-3)], case="rbc")
    _compare(big, 1, 3)
    # second call: a smaller plan, on the scratch array the first call left filled
    small = RBCBatch((64, 48), [dict(ra=1e5, pr=1.0, dt=0.01), dict(ra=3e5, pr=1.0, dt=0.005)],
                     case="rbc", integrator="rk3")
    _compare(small, 2, 2)


def test_noalloc_step_rejects_a_plan_beyond_the_scratch_size():
    import torch
    from pypde_amd._lib import lib
    from pypde_amd.engine import RBCBatch
    eng = RBCBatch((256, 256), [dict(ra=1e8, pr=1.0, dt=1e-3)] * 4, case="rbc")
    h = eng.build_plan().handle
    assert lib.pypde_plan_region_bytes(h, 3) > lib.pypde_scratch_bytes()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    X0 = eng.X.clone()
    assert lib.pypde_rbc_step_noalloc(h, vp(eng.X), vp(eng.pres), vp(eng.P), 1, s) != 0
    assert b"scratch" in lib.pypde_last_error()
    torch.cuda.synchronize()
    assert torch.equal(eng.X, X0)
