"""HDF5 state files in the reference's layout (pypde/field.py:83-173,
navier/rbc2d_base.py:205-257), through pypde_amd/libpypde_h5.so (the HDF5 C
library; include/pypde_amd_h5.h).

    with H5File("flow.h5", "a") as hf:
        hf.write("temp/vhat", array)       # create, or overwrite in place
        hf.read("temp/vhat")               # numpy float64, None if missing
        hf.keys()                          # every link name, like hf.visit
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpypde_h5.so")

READ, APPEND = 0, 1
MISSING = 1

_lib = None


def lib():
    """The loaded state-file library (ImportError when it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("pypde_amd HDF5 library missing (%s): run python pypde_amd/build.py "
                              "with the HDF5 C library available" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        i64, vp, cp = ctypes.c_int64, ctypes.c_void_p, ctypes.c_char_p
        L.pypde_h5_open.argtypes, L.pypde_h5_open.restype = [cp, ctypes.c_int], i64
        L.pypde_h5_close.argtypes, L.pypde_h5_close.restype = [i64], ctypes.c_int
        L.pypde_h5_write.argtypes = [i64, cp, vp, ctypes.c_int, vp]
        L.pypde_h5_write.restype = ctypes.c_int
        L.pypde_h5_shape.argtypes = [i64, cp, ctypes.POINTER(ctypes.c_int), vp]
        L.pypde_h5_shape.restype = ctypes.c_int
        L.pypde_h5_read.argtypes, L.pypde_h5_read.restype = [i64, cp, vp, i64], ctypes.c_int
        L.pypde_h5_list.argtypes, L.pypde_h5_list.restype = [i64, vp, i64], i64
        L.pypde_h5_last_error.argtypes, L.pypde_h5_last_error.restype = [], cp
        _lib = L
    return _lib


def exported_symbols():
    return ["pypde_h5_open", "pypde_h5_close", "pypde_h5_write", "pypde_h5_shape",
            "pypde_h5_read", "pypde_h5_list", "pypde_h5_last_error"]


def _check(rc, what):
    if rc < 0:
        raise IOError("%s: %s" % (what, lib().pypde_h5_last_error().decode()))
    return rc


def _host(data):
    """float64, C-contiguous host copy of a scalar, array or tensor."""
    if hasattr(data, "detach"):
        data = data.detach().cpu().numpy()
    # np.array keeps 0-d values 0-d (a scalar dataspace); ascontiguousarray would not
    return np.array(data, dtype=np.float64, order="C")


class H5File:
    """mode "a": read/write, created if missing; mode "r": read only (h5py's modes)."""

    def __init__(self, filename, mode="a"):
        if mode not in ("a", "r"):
            raise ValueError("mode must be 'a' or 'r'")
        self.filename = str(filename)
        h = lib().pypde_h5_open(self.filename.encode(), APPEND if mode == "a" else READ)
        self.handle = _check(h, "open " + self.filename)

    def close(self):
        if self.handle is not None:
            _check(lib().pypde_h5_close(self.handle), "close " + self.filename)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __contains__(self, name):
        return self.shape(name) is not None

    def write(self, name, data):
        a = _host(data)
        dims = np.array(a.shape, dtype=np.int64)
        _check(lib().pypde_h5_write(self.handle, name.encode(), a.ctypes.data, a.ndim,
                                    dims.ctypes.data if a.ndim else None), "write " + name)

    def shape(self, name):
        nd = ctypes.c_int(0)
        dims = np.zeros(8, dtype=np.int64)
        rc = _check(lib().pypde_h5_shape(self.handle, name.encode(), ctypes.byref(nd),
                                         dims.ctypes.data), "shape " + name)
        if rc == MISSING:
            return None
        return tuple(int(d) for d in dims[:nd.value])

    def read(self, name):
        shape = self.shape(name)
        if shape is None:
            return None
        out = np.empty(shape, dtype=np.float64)
        _check(lib().pypde_h5_read(self.handle, name.encode(), out.ctypes.data, out.size),
               "read " + name)
        return out

    def keys(self):
        n = _check(lib().pypde_h5_list(self.handle, None, 0), "list")
        buf = ctypes.create_string_buffer(int(n))
        _check(lib().pypde_h5_list(self.handle, buf, n), "list")
        return [k for k in buf.value.decode().split("\n") if k]
