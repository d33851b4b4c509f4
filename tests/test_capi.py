"""The C-ABI library loads and exports every symbol include/pypde_amd.h declares
(no kernel is launched: CPU only)."""
import ctypes
import os
import re

import pytest

from conftest import REPO


def declared_symbols(header="pypde_amd.h"):
    src = open(os.path.join(REPO, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted({m.group(1) for m in re.finditer(r"\b(pypde_[a-z0-9_]+)\s*\(", src)})


def test_every_header_is_checked():
    assert sorted(f for f in os.listdir(os.path.join(REPO, "include")) if f.endswith(".h")) == \
        ["pypde_amd.h", "pypde_amd_h5.h"]


def test_library_exports_every_declared_symbol():
    from pypde_amd import _lib, h5io
    for header, path in (("pypde_amd.h", _lib.LIB_PATH), ("pypde_amd_h5.h", h5io.LIB_PATH)):
        lib = ctypes.CDLL(path)
        names = declared_symbols(header)
        assert len(names) >= 7, header
        missing = [n for n in names if not hasattr(lib, n)]
        assert not missing, (header, missing)


def test_version_and_errors_without_gpu():
    from pypde_amd import _lib
    assert _lib.lib.pypde_version() >= 1
    # argument validation happens before any HIP call
    g = _lib.Lines()
    g.axis = 3
    g.src = 1
    g.dst = 1
    g.len_in = g.len_out = 8
    rc = _lib.lib.pypde_fdma_solve(ctypes.byref(g), 1, 1, 1, 1, 0, None)
    assert rc != 0
    assert b"axis" in _lib.lib.pypde_last_error()


def test_every_declared_symbol_has_a_python_signature():
    from pypde_amd import _lib
    assert set(declared_symbols()) <= set(_lib.exported_symbols())


def test_step_plan_host_logic_without_gpu():
    """Step plans are host objects: creation checks its regions, recording
    copies descriptors, and the sizes read back (no kernel is launched)."""
    from pypde_amd import _lib
    lib = _lib.lib
    h = ctypes.c_void_p()
    bases = (ctypes.c_void_p * 2)(4096, 8192)
    sizes = (ctypes.c_longlong * 2)(4096, 100)
    assert lib.pypde_plan_create(2, bases, sizes, ctypes.byref(h)) == 0
    assert lib.pypde_plan_size(h) == 0
    assert lib.pypde_plan_region_bytes(h, 1) == 100
    assert lib.pypde_plan_region_bytes(h, 2) == -1
    d = _lib.GemmDesc()
    assert lib.pypde_plan_add_gemm(h, ctypes.byref(d)) == 0
    assert lib.pypde_plan_add_grouped(h, ctypes.byref(d), 5) != 0
    assert b"1..4" in lib.pypde_last_error()
    assert lib.pypde_plan_size(h) == 1
    # zero steps run nothing; an RBC step needs the four regions
    assert lib.pypde_plan_run(h, None, 0, None) == 0
    assert lib.pypde_rbc_step(h, 1, 1, 1, 1, 1, None) != 0
    assert b"regions" in lib.pypde_last_error()
    assert lib.pypde_plan_destroy(h) == 0
    # overlapping regions are rejected
    sizes = (ctypes.c_longlong * 2)(4097, 100)
    h2 = ctypes.c_void_p()
    assert lib.pypde_plan_create(2, bases, sizes, ctypes.byref(h2)) != 0
    assert b"overlap" in lib.pypde_last_error()
    assert not h2.value


def test_engine_launch_records_without_gpu():
    """Every kind of launch the fused step issues can be recorded into a step
    plan (host only: recording copies the arguments, nothing runs)."""
    from pypde_amd import _lib
    from pypde_amd.engine import Plan, _Call, _Gemm, _GemmGroup, _RecCall, _lines, _rec
    base = 1 << 20
    bases = (ctypes.c_void_p * 4)(base, 2 * base, 3 * base, 4 * base)
    sizes = (ctypes.c_longlong * 4)(base, base, base, base)
    plan = Plan(4, bases, sizes)
    lib = _lib.lib
    g = _Gemm("g", m=4, n=4, k=4, batch=1, A=base, lda=4, B=base, ldb=4, C=base, ldc=4)
    plan.add(g)
    plan.add(_GemmGroup([g, _Gemm("h", m=4, n=4, k=4, batch=1, A=base, lda=4, B=base, ldb=4,
                                  C=2 * base, ldc=4)]))
    gl = _lines(base, base, 4, 16, 4, 4, 1, 0)
    plan.add(_RecCall("r", gl, {v: _rec(3 * base, 4, 16, v) for v in (1, 2, 3)}, 2))
    vp = ctypes.c_void_p
    plan.add(_Call("p", lib.pypde_rbc_products, vp(base), vp(base), vp(base), 6, 6, 1))
    plan.add(_Call("py", lib.pypde_rbc_products_yfold, vp(base), None, vp(base), vp(base), 6, 6,
                   1.0, -1.0, 0.0, 1))
    plan.add(_Call("s", lib.pypde_splitk_reduce, vp(base), 2, vp(base), None, 1, 0, vp(base),
                   16, 1))
    a = _lib.RbcRhs()
    plan.add(_Call("b", lib.pypde_rbc_rhs_banded, ctypes.byref(a)))
    assert plan.size() == 7
    # the adjoint's launches, as ops records them
    from pypde_amd import ops
    plan.add(ops.Recorded("gemm", "g", descs=[g.desc]))
    plan.add(ops.Recorded("group", "gg", descs=[g.desc, g.desc]))
    for fn, args in ((lib.pypde_axpby, (16, 1.0, vp(base), 0.0, vp(2 * base))),
                     (lib.pypde_transpose, (vp(base), 4, 16, vp(2 * base), 4, 16, 4, 4, 1)),
                     (lib.pypde_scaled_sub, (16, 1, vp(base), 0.5, vp(base), vp(2 * base))),
                     (lib.pypde_set_entries, (vp(base), 16, 1, 0.0)),
                     (lib.pypde_rbc_products_vjp, (vp(base), vp(base), vp(base), vp(2 * base),
                                                   6, 6, 1)),
                     (lib.pypde_rbc_products_vjp_c, (vp(base), vp(base), vp(base), vp(base),
                                                     vp(2 * base), vp(3 * base), 0.5, 6, 6, 1)),
                     (lib.pypde_recurrence_solve, (ctypes.byref(gl),
                                                   ctypes.byref(_rec(3 * base, 4, 16, 2))))):
        plan.add(ops.Recorded("call", fn.__name__, fn=fn, args=args))
    assert plan.size() == 16
    with pytest.raises(TypeError):
        plan.add(ops.Recorded("call", "x", fn=lib.pypde_version, args=()))
    # a VJP plan carves its workspace regions (4 on) from one buffer, 256-byte aligned
    assert lib.pypde_rbc_step_vjp_workspace_bytes(plan.handle) == -1
    bases6 = (ctypes.c_void_p * 6)(*[base * (i + 1) for i in range(6)])
    sizes6 = (ctypes.c_longlong * 6)(base, base, base, base, 100, 300)
    vplan = Plan(6, bases6, sizes6)
    assert lib.pypde_rbc_step_vjp_workspace_bytes(vplan.handle) == 256 + 512
    assert lib.pypde_rbc_step_vjp(vplan.handle, 1, 1, 1, 1, 255, None) != 0
    assert b"aligned" in lib.pypde_last_error()
    # a plan of four regions has no workspace region: rejected before anything runs
    assert lib.pypde_rbc_step_vjp(plan.handle, 1, 1, 1, 1, 256, None) != 0
    assert b"regions" in lib.pypde_last_error()


def test_multi_launch_packing_without_gpu():
    """The multi launches' host packing (no kernel runs): member ranges, the
    work total and the argument checks."""
    from pypde_amd import _lib
    lib = _lib.lib
    info = _lib.MultiInfo()
    nb = lib.pypde_multi_table_bytes(1)
    assert nb > 0 and lib.pypde_multi_table_bytes(9) == -1
    host = (ctypes.c_ubyte * nb)()
    descs = (_lib.GemmDesc * 3)()
    for i, (m, n, b) in enumerate(((64, 64, 2), (0, 64, 1), (33, 65, 1))):
        d = descs[i]
        d.m, d.n, d.k, d.batch = m, n, 16, b
        d.A = d.B = d.C = 4096
        d.lda = d.ldb = d.ldc = 128
    # 32 x 32 tiles: 2*2*2 + 0 + 2*3*1 = 14
    assert lib.pypde_gemm_multi_pack(descs, 3, 2, host, nb, ctypes.byref(info)) == 0
    assert (info.kind, info.count, info.variant, info.total) == (1, 3, 2, 14)
    # a split-K variant has no multi form; a folded member is rejected
    assert lib.pypde_gemm_multi_pack(descs, 3, 9, host, nb, ctypes.byref(info)) != 0
    descs[2].fold = 1
    assert lib.pypde_gemm_multi_pack(descs, 3, 2, host, nb, ctypes.byref(info)) != 0
    assert b"folded" in lib.pypde_last_error()
    # recurrences: lines per member, chunk code from the longest line
    g = (_lib.Lines * 2)()
    r = (_lib.Recurrence * 2)()
    for i, (n, nl, b) in enumerate(((190, 64, 3), (62, 30, 2))):
        g[i].src = g[i].dst = 4096
        g[i].len_in = g[i].len_out = n
        g[i].nlines, g[i].batch, g[i].axis = nl, b, 0
        g[i].src_ld = g[i].dst_ld = nl
        r[i].LM = r[i].R = r[i].MU = r[i].NU = 8192
    nb = lib.pypde_multi_table_bytes(2)
    host = (ctypes.c_ubyte * nb)()
    assert lib.pypde_recurrence_multi_pack(g, r, 2, host, nb, ctypes.byref(info)) == 0
    assert info.total == 64 * 3 + 30 * 2 and info.flags == (0 | (0 << 1) | (2 << 2))
