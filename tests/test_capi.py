"""The C-ABI library loads and exports every symbol include/pypde_amd.h declares
(no kernel is launched: CPU only)."""
import ctypes
import os
import re

from conftest import REPO


def declared_symbols():
    hdrs = [os.path.join(REPO, "include", f) for f in os.listdir(os.path.join(REPO, "include"))
            if f.endswith(".h")]
    names = set()
    for h in hdrs:
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b(pypde_[a-z0-9_]+)\s*\(", src):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from pypde_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


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
