"""ctypes binding of libpypde_hip.so (the C-ABI in include/pypde_amd.h).

The product path has no CPU fallback: importing this module without the built
library, or calling a kernel without a visible GPU, raises.
"""
import ctypes
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpypde_hip.so")


class PypdeError(RuntimeError):
    pass


class Lines(ctypes.Structure):
    """Mirror of `pypde_lines` (include/pypde_amd.h)."""

    _fields_ = [
        ("src", ctypes.c_void_p),
        ("dst", ctypes.c_void_p),
        ("src_ld", ctypes.c_longlong),
        ("dst_ld", ctypes.c_longlong),
        ("src_bs", ctypes.c_longlong),
        ("dst_bs", ctypes.c_longlong),
        ("len_in", ctypes.c_int),
        ("len_out", ctypes.c_int),
        ("nlines", ctypes.c_int),
        ("batch", ctypes.c_int),
        ("axis", ctypes.c_int),
        ("lines_per_block", ctypes.c_int),
    ]


class BIdx(ctypes.Structure):
    """Mirror of `pypde_bidx`."""

    _fields_ = [("s1", ctypes.c_longlong), ("s2", ctypes.c_longlong), ("d1", ctypes.c_int),
                ("m1", ctypes.c_int), ("d2", ctypes.c_int), ("m2", ctypes.c_int)]


class GemmDesc(ctypes.Structure):
    """Mirror of `pypde_gemm_desc`."""

    _fields_ = [("transA", ctypes.c_int), ("transB", ctypes.c_int), ("m", ctypes.c_int),
                ("n", ctypes.c_int), ("k", ctypes.c_int), ("batch", ctypes.c_int),
                ("alpha", ctypes.c_double), ("beta", ctypes.c_double),
                ("A", ctypes.c_void_p), ("B", ctypes.c_void_p), ("D", ctypes.c_void_p),
                ("C", ctypes.c_void_p), ("lda", ctypes.c_longlong), ("ldb", ctypes.c_longlong),
                ("ldd", ctypes.c_longlong), ("ldc", ctypes.c_longlong), ("iA", BIdx), ("iB", BIdx),
                ("iD", BIdx), ("iC", BIdx), ("alpha_vec", ctypes.c_void_p),
                ("alpha_div", ctypes.c_int), ("variant", ctypes.c_int),
                ("E", ctypes.c_void_p), ("lde", ctypes.c_longlong), ("iE", BIdx),
                ("gamma_vec", ctypes.c_void_p), ("gamma", ctypes.c_double),
                ("gamma_div", ctypes.c_int), ("zero_first", ctypes.c_int),
                ("csn", ctypes.c_longlong), ("fold", ctypes.c_int), ("A1", ctypes.c_void_p),
                ("B1", ctypes.c_void_p), ("lda1", ctypes.c_longlong), ("ldb1", ctypes.c_longlong),
                ("k1", ctypes.c_int), ("ska", ctypes.c_longlong), ("skb", ctypes.c_longlong),
                ("fold_len", ctypes.c_int), ("fold_sign", ctypes.c_uint),
                ("mirror_sign", ctypes.c_double), ("mirror_div", ctypes.c_int)]


class RbcRhs(ctypes.Structure):
    """Mirror of `pypde_rbc_rhs`."""

    _fields_ = [("X", ctypes.c_void_p), ("pres", ctypes.c_void_p), ("cst", ctypes.c_void_p),
                ("R", ctypes.c_void_p), ("mdta", ctypes.c_void_p), ("dta", ctypes.c_void_p),
                ("fdiff", ctypes.c_void_p), ("tx", ctypes.c_void_p), ("ty", ctypes.c_void_p),
                ("M0", ctypes.c_int), ("M1", ctypes.c_int), ("N0", ctypes.c_int),
                ("N1", ctypes.c_int), ("batch", ctypes.c_int), ("rows_per_thread", ctypes.c_int)]


class MultiInfo(ctypes.Structure):
    """Mirror of `pypde_multi_info`."""

    _fields_ = [("kind", ctypes.c_int), ("count", ctypes.c_int), ("variant", ctypes.c_int),
                ("flags", ctypes.c_int), ("total", ctypes.c_longlong)]


class Recurrence(ctypes.Structure):
    """Mirror of `pypde_recurrence`."""

    _fields_ = [("F", ctypes.c_void_p), ("LM", ctypes.c_void_p), ("R", ctypes.c_void_p),
                ("MU", ctypes.c_void_p), ("NU", ctypes.c_void_p), ("pre", ctypes.c_void_p),
                ("entry_stride", ctypes.c_longlong), ("per_line", ctypes.c_int),
                ("variant", ctypes.c_int)]


if not os.path.exists(LIB_PATH):
    raise ImportError(
        "pypde_amd: %s is missing -- build the HIP library first "
        "(python pypde_amd/build.py or __graft_entry__.build())" % LIB_PATH)

lib = ctypes.CDLL(LIB_PATH)

_vp = ctypes.c_void_p
_d = ctypes.c_double
_i = ctypes.c_int
_ll = ctypes.c_longlong
_lp = ctypes.POINTER(Lines)

_SIGS = {
    "pypde_version": ([], _i),
    "pypde_last_error": ([], ctypes.c_char_p),
    "pypde_device_count": ([], _i),
    "pypde_dgemm": ([_i, _i, _i, _i, _i, _d, _vp, _ll, _ll, _vp, _ll, _ll, _d, _vp, _ll, _ll,
                     _vp, _ll, _ll, _i, _vp], _i),
    "pypde_stencil_apply": ([_lp, _vp, _vp], _i),
    "pypde_stencil_inverse": ([_lp, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "pypde_cheb_deriv": ([_lp, _vp, _i, _d, _vp], _i),
    "pypde_banded_apply": ([_lp, _i, ctypes.POINTER(_i), _vp, _vp], _i),
    "pypde_fdma_solve": ([_lp, _vp, _vp, _vp, _vp, _ll, _vp], _i),
    "pypde_fdma_solve_t": ([_lp, _vp, _vp, _vp, _vp, _ll, _vp], _i),
    "pypde_poisson_solve": ([_lp, _vp, _vp, _i, _vp], _i),
    "pypde_twodma_solve": ([_lp, _vp, _vp, _vp], _i),
    "pypde_pdma_solve": ([_lp, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "pypde_axpby": ([_ll, _d, _vp, _d, _vp, _vp], _i),
    "pypde_dgemm_batched": ([ctypes.c_void_p, _vp], _i),
    "pypde_dgemm_grouped": ([ctypes.c_void_p, _i, _vp], _i),
    "pypde_splitk_reduce": ([_vp, _i, _vp, _vp, _i, _i, _vp, ctypes.c_longlong, _i, _vp], _i),
    "pypde_rbc_products": ([_vp, _vp, _vp, _i, _i, _i, _vp], _i),
    "pypde_rbc_products_yfold": ([_vp, _vp, _vp, _vp, _i, _i, _d, _d, _d, _i, _vp], _i),
    "pypde_scaled_sub": ([_ll, _i, _vp, _d, _vp, _vp, _vp], _i),
    "pypde_set_entries": ([_vp, _ll, _i, _d, _vp], _i),
    "pypde_recurrence_solve": ([_lp, ctypes.c_void_p, _vp], _i),
    "pypde_rbc_products_vjp": ([_vp, _vp, _vp, _vp, _i, _i, _i, _vp], _i),
    "pypde_rbc_products_vjp_c": ([_vp, _vp, _vp, _vp, _vp, _vp, _d, _i, _i, _i, _vp], _i),
    "pypde_plan_create": ([_i, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p], _i),
    "pypde_plan_destroy": ([_vp], _i),
    "pypde_plan_add_gemm": ([_vp, ctypes.c_void_p], _i),
    "pypde_plan_add_grouped": ([_vp, ctypes.c_void_p, _i], _i),
    "pypde_plan_add_recurrence": ([_vp, _lp, ctypes.c_void_p], _i),
    "pypde_plan_add_products": ([_vp, _vp, _vp, _vp, _i, _i, _i], _i),
    "pypde_plan_add_products_yfold": ([_vp, _vp, _vp, _vp, _vp, _i, _i, _d, _d, _d, _i], _i),
    "pypde_plan_add_splitk_reduce": ([_vp, _vp, _i, _vp, _vp, _i, _i, _vp, ctypes.c_longlong,
                                      _i], _i),
    "pypde_plan_add_rhs_banded": ([_vp, ctypes.c_void_p], _i),
    "pypde_plan_add_fdma": ([_vp, _lp, _vp, _vp, _vp, _vp, ctypes.c_longlong], _i),
    "pypde_plan_add_poisson": ([_vp, _lp, _vp, _vp, _i], _i),
    "pypde_rbc_rhs_banded": ([ctypes.c_void_p, _vp], _i),
    "pypde_rbc_products_rhs": ([_vp, _vp, _vp, _vp, _i, _i, _d, _d, _d, _i, ctypes.c_void_p,
                                _vp], _i),
    "pypde_plan_add_products_rhs": ([_vp, _vp, _vp, _vp, _vp, _i, _i, _d, _d, _d, _i,
                                     ctypes.c_void_p], _i),
    "pypde_multi_table_bytes": ([_i], ctypes.c_longlong),
    "pypde_gemm_multi_pack": ([ctypes.c_void_p, _i, _i, _vp, ctypes.c_longlong, ctypes.c_void_p],
                              _i),
    "pypde_dgemm_multi": ([_vp, ctypes.c_void_p, _vp], _i),
    "pypde_recurrence_multi_pack": ([ctypes.c_void_p, ctypes.c_void_p, _i, _vp,
                                     ctypes.c_longlong, ctypes.c_void_p], _i),
    "pypde_recurrence_multi": ([_vp, ctypes.c_void_p, _vp], _i),
    "pypde_products_multi_pack": ([ctypes.c_void_p] * 10 + [_i, _vp, ctypes.c_longlong,
                                                              ctypes.c_void_p], _i),
    "pypde_products_multi": ([_vp, ctypes.c_void_p, _vp], _i),
    "pypde_rhs_multi_pack": ([ctypes.c_void_p, _i, _vp, ctypes.c_longlong, ctypes.c_void_p], _i),
    "pypde_rhs_multi": ([_vp, ctypes.c_void_p, _vp], _i),
    "pypde_cheb_dct": ([_lp, _i, _vp, _vp, _vp, _i, _vp], _i),
    "pypde_plan_add_multi": ([_vp, _vp, ctypes.c_void_p], _i),
    "pypde_rbc_step_batched": ([_vp, _i, _vp], _i),
    "pypde_transpose": ([_vp, _ll, _ll, _vp, _ll, _ll, _i, _i, _i, _vp], _i),
    "pypde_plan_add_axpby": ([_vp, _ll, _d, _vp, _d, _vp], _i),
    "pypde_plan_add_transpose": ([_vp, _vp, _ll, _ll, _vp, _ll, _ll, _i, _i, _i], _i),
    "pypde_plan_add_scaled_sub": ([_vp, _ll, _i, _vp, _d, _vp, _vp], _i),
    "pypde_plan_add_set_entries": ([_vp, _vp, _ll, _i, _d], _i),
    "pypde_plan_add_products_vjp": ([_vp, _vp, _vp, _vp, _vp, _i, _i, _i], _i),
    "pypde_plan_add_products_vjp_c": ([_vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _i, _i, _i], _i),
    "pypde_rbc_step_vjp_workspace_bytes": ([_vp], _ll),
    "pypde_rbc_step_vjp": ([_vp, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "pypde_plan_size": ([_vp], _i),
    "pypde_plan_region_bytes": ([_vp, _i], ctypes.c_longlong),
    "pypde_plan_run": ([_vp, ctypes.c_void_p, _i, _vp], _i),
    "pypde_rbc_step": ([_vp, _vp, _vp, _vp, _vp, _i, _vp], _i),
    "pypde_scratch_bytes": ([], ctypes.c_longlong),
    "pypde_rbc_step_noalloc": ([_vp, _vp, _vp, _vp, _i, _vp], _i),
}

for _name, (_args, _res) in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.argtypes = _args
    _fn.restype = _res


def exported_symbols():
    return sorted(_SIGS)


def check(rc, what=""):
    if rc != 0:
        msg = lib.pypde_last_error()
        raise PypdeError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))


_gpu_checked = False


def require_gpu():
    """Raise unless a HIP device is visible; the kernels have no CPU path."""
    global _gpu_checked
    if not _gpu_checked:
        if not torch.cuda.is_available():
            raise PypdeError("pypde_amd needs a HIP (ROCm) GPU; none is visible")
        _gpu_checked = True


def stream_ptr():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None
