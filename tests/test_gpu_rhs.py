"""The banded Helmholtz right-hand side of the fused step (the sparse
solve_rhs / solve_old plan applies of templates/hholtz.py:70-93 on the
pressure gradient, old value, buoyancy and explicit diffusion,
navier/rbc2d.py:213-283): every kernel form -- one output per thread, row
strips of 4 and 8, LDS tiles of 16 x 64 -- alone and as a multi launch, and
the launch merged with the products, give the same bits as the one-output form
(which the golden steps pin against the reference)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def host(t):
    return t.detach().cpu().numpy()


def _rhs_struct(eng):
    from pypde_amd._lib import RbcRhs
    for l in eng.launches:
        for a in getattr(l, "args", ()):
            obj = getattr(a, "_obj", None)
            if isinstance(obj, RbcRhs):
                return RbcRhs.from_buffer_copy(obj)
    raise AssertionError("no banded right-hand side in the step")


@pytest.mark.parametrize("shape,beta,integrator", [((20, 16), 1.0, "eu"), ((70, 100), 0.5, "eu"),
                                                   ((256, 256), 1.0, "eu"),
                                                   ((96, 64), 1.0, "rk3")])
def test_rhs_forms_bit_identical(shape, beta, integrator):
    import torch
    from pypde_amd._lib import MultiInfo, lib
    from pypde_amd.engine import RBCBatch
    params = [dict(ra=1e5, pr=1.0, dt=0.01), dict(ra=1e6, pr=0.7, dt=0.005)]
    eng = RBCBatch(shape, params, case="rbc", beta=beta, integrator=integrator)
    rng = np.random.default_rng(3)
    dev = eng.X.device
    eng.X.copy_(torch.as_tensor(rng.standard_normal(tuple(eng.X.shape)), device=dev))
    eng.pres.copy_(torch.as_tensor(rng.standard_normal(tuple(eng.pres.shape)), device=dev))
    a = _rhs_struct(eng)
    R = torch.empty((eng.B, 3, eng.M0, eng.M1), dtype=torch.float64, device=dev)
    a.R = R.data_ptr()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = {}
    for rs in (1, 4, 8, 16):
        a.rows_per_thread = rs
        R.fill_(float("nan"))
        assert lib.pypde_rbc_rhs_banded(ctypes.byref(a), s) == 0
        torch.cuda.synchronize()
        out[("single", rs)] = host(R).copy()
        nb = lib.pypde_multi_table_bytes(4)
        buf = (ctypes.c_ubyte * nb)()
        info = MultiInfo()
        assert lib.pypde_rhs_multi_pack(ctypes.byref(a), 1, buf, nb, ctypes.byref(info)) == 0
        table = torch.frombuffer(bytearray(buf), dtype=torch.uint8).to(dev)
        R.fill_(float("nan"))
        assert lib.pypde_rhs_multi(ctypes.c_void_p(table.data_ptr()), ctypes.byref(info), s) == 0
        torch.cuda.synchronize()
        out[("multi", rs)] = host(R).copy()
    ref = out[("single", 1)]
    assert np.isfinite(ref).all()
    for k, v in out.items():
        assert np.array_equal(v, ref), k
    # the launch merged with the y-folded products (when the layout is folded)
    if eng.yfold is not None:
        a.rows_per_thread = 1
        R.fill_(float("nan"))
        L0, L1 = eng.L0, eng.L1
        Y = torch.zeros((eng.B, 3, 2 * L0, 2 * L1), dtype=torch.float64, device=dev)
        P = torch.empty((eng.B, 3, L0, L1), dtype=torch.float64, device=dev)
        assert lib.pypde_rbc_products_rhs(ctypes.c_void_p(Y.data_ptr()), None,
                                          ctypes.c_void_p(eng.dTbcdz1.data_ptr()),
                                          ctypes.c_void_p(P.data_ptr()), L0, L1, 1.0, 1.0, 0.0,
                                          eng.B, ctypes.byref(a), s) == 0
        torch.cuda.synchronize()
        assert np.array_equal(host(R), ref)
        a.rows_per_thread = 4
        assert lib.pypde_rbc_products_rhs(ctypes.c_void_p(Y.data_ptr()), None,
                                          ctypes.c_void_p(eng.dTbcdz1.data_ptr()),
                                          ctypes.c_void_p(P.data_ptr()), L0, L1, 1.0, 1.0, 0.0,
                                          eng.B, ctypes.byref(a), s) != 0
        assert b"rows_per_thread" in lib.pypde_last_error()
