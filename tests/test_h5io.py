"""HDF5 state files (pypde_amd/h5io.py, libpypde_h5.so) on the host.

The reference writes through h5py (pypde/field.py:83-173), which is absent
here, so no file written by the reference exists to compare with; the layout is
pinned by what h5py's create_dataset makes (contiguous little-endian float64,
scalar dataspace for 0-d values, intermediate groups) as read back by the HDF5
tools' own reader (h5dump), and the reader by a file the HDF5 tools' own writer
made (h5import)."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest
import torch

from pypde_amd import h5io
from pypde_amd.field import FieldBase

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
H5IMPORT = shutil.which("h5import") or "/opt/conda/bin/h5import"
tools = pytest.mark.skipif(not (os.path.exists(H5DUMP) and os.path.exists(H5IMPORT)),
                           reason="HDF5 command-line tools not installed")


def test_library_exports_every_declared_symbol():
    lib = h5io.lib()
    hdr = open(os.path.join(os.path.dirname(h5io.__file__), "..", "include",
                            "pypde_amd_h5.h")).read()
    declared = set(re.findall(r"\b(pypde_h5_\w+)\s*\(", hdr))
    assert declared == set(h5io.exported_symbols())
    for name in declared:
        assert hasattr(lib, name), name


def test_round_trip_ranks_groups_and_overwrite(tmp_path):
    fn = str(tmp_path / "a.h5")
    rng = np.random.default_rng(0)
    arrs = {"s": np.float64(2.5), "one": rng.standard_normal(7),
            "temp/v": rng.standard_normal((5, 6)), "a/b/c": rng.standard_normal((2, 3, 4))}
    with h5io.H5File(fn, "a") as hf:
        for k, v in arrs.items():
            hf.write(k, v)
    with h5io.H5File(fn, "r") as hf:
        for k, v in arrs.items():
            got = hf.read(k)
            assert got.shape == np.shape(v)
            assert np.array_equal(got, v)
        assert hf.read("missing") is None
        assert hf.read("temp/missing") is None
        assert "temp/v" in hf and "temp/w" not in hf
        assert hf.keys() == ["a", "a/b", "a/b/c", "one", "s", "temp", "temp/v"]
    # mode "a" re-opens and overwrites in place; another shape is an error
    with h5io.H5File(fn, "a") as hf:
        hf.write("temp/v", -arrs["temp/v"])
        with pytest.raises(IOError):
            hf.write("temp/v", np.zeros((6, 5)))
        hf.write("s", 3.0)
    with h5io.H5File(fn, "r") as hf:
        assert np.array_equal(hf.read("temp/v"), -arrs["temp/v"])
        assert float(hf.read("s")) == 3.0
    with pytest.raises(IOError):
        h5io.H5File(str(tmp_path / "nope.h5"), "r")
    with pytest.raises(ValueError):
        h5io.H5File(fn, "w")


class _HostField(FieldBase):
    """The Field IO methods on host tensors (Field itself lives on the GPU)."""

    def __init__(self, n0, n1, m0, m1, seed):
        g = torch.Generator().manual_seed(seed)
        self.v = torch.randn(n0, n1, generator=g, dtype=torch.float64)
        self.vhat = torch.randn(m0, m1, generator=g, dtype=torch.float64)
        self.t = 0.25
        self._x = np.linspace(-1, 1, n0)
        self._y = np.linspace(-1, 1, n1)

    @property
    def x(self):
        return self._x

    @property
    def y(self):
        return self._y


def dump(path, *args):
    r = subprocess.run([H5DUMP] + list(args) + [path], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return r.stdout


@tools
def test_field_layout_read_by_h5dump(tmp_path):
    """Field.write in the reference's layout; h5dump sees h5py's default objects
    and the values in row-major order."""
    fn = str(tmp_path / "flow.h5")
    f = _HostField(6, 5, 4, 3, seed=1)
    f.write(filename=fn, dict={"ra": 1e5, "pr": 1.0}, grp_name="temp")
    txt = dump(fn, "-p", "-H")
    for name in ("v", "vhat"):
        assert re.search(r'DATASET "%s" \{\s*DATATYPE\s+H5T_IEEE_F64LE' % name, txt)
    assert 'GROUP "temp"' in txt
    assert "DATASPACE  SIMPLE { ( 6, 5 ) / ( 6, 5 ) }" in txt
    assert "DATASPACE  SIMPLE { ( 4, 3 ) / ( 4, 3 ) }" in txt
    for name in ("time", "ra", "pr"):
        assert re.search(r'DATASET "%s" \{\s*DATATYPE\s+H5T_IEEE_F64LE\s*DATASPACE\s+SCALAR'
                         % name, txt), name
    assert "CHUNKED" not in txt and txt.count("CONTIGUOUS") == 7
    # values, row by row, as h5dump prints them
    out = dump(fn, "-d", "temp/vhat", "-m", "%.17g", "-w", "0")
    rows = re.findall(r"\((\d+),0\):([^\n]*)", out)
    got = np.array([[float(x) for x in r.replace(",", " ").split()] for _, r in rows])
    assert np.array_equal(got, f.vhat.numpy())
    # read back through Field.read, in place
    g = _HostField(6, 5, 4, 3, seed=2)
    ptr = g.vhat.data_ptr()
    meta = {"ra": 0.0, "pr": 0.0}
    g.read(filename=fn, dict=meta, grp_name="temp")
    assert g.vhat.data_ptr() == ptr
    assert torch.equal(g.vhat, f.vhat) and torch.equal(g.v, f.v)
    assert g.t == 0.25 and float(meta["ra"]) == 1e5


@tools
def test_reader_on_a_file_from_h5import(tmp_path):
    data = np.arange(12.0).reshape(3, 4) + 0.5
    txt = tmp_path / "in.txt"
    np.savetxt(txt, data)
    cfg = tmp_path / "cfg.txt"
    cfg.write_text("PATH temp/vhat\nINPUT-CLASS TEXTFP\nRANK 2\nDIMENSION-SIZES 3 4\n"
                   "OUTPUT-CLASS FP\nOUTPUT-SIZE 64\nOUTPUT-ARCHITECTURE IEEE\n"
                   "OUTPUT-BYTE-ORDER LE\n")
    fn = str(tmp_path / "imp.h5")
    r = subprocess.run([H5IMPORT, str(txt), "-c", str(cfg), "-o", fn], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    with h5io.H5File(fn, "r") as hf:
        assert hf.shape("temp/vhat") == (3, 4)
        assert np.array_equal(hf.read("temp/vhat"), data)


def test_missing_dataset_reads_zero_and_lists_keys(tmp_path, capsys):
    """read_single_hdf5 (field.py:135-144): prints the keys, returns 0.0."""
    fn = str(tmp_path / "m.h5")
    with h5io.H5File(fn, "a") as hf:
        hf.write("x", np.arange(3.0))
    with h5io.H5File(fn, "r") as hf:
        assert FieldBase.read_single_hdf5(hf, "time") == 0.0
    out = capsys.readouterr().out
    assert "Cannot read time" in out and "Group: x" in out


def test_filename_rule():
    """field.py:91-97: leading_str + _{t:07.2f} + .h5 when no filename is given."""
    assert FieldBase._filename(None, "flow", True, 12.3456) == "flow_0012.35.h5"
    assert FieldBase._filename(None, "flow", False, 1.0) == "flow.h5"
    assert FieldBase._filename("given.h5", "flow", True, 1.0) == "given.h5"
