"""The C-ABI library loads and exports every symbol include/pypde_amd.h declares
(no kernel is launched: CPU only)."""
import ctypes
import os
import re

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
