"""NavierStokes state files on the GPU: write -> read into a fresh solver gives
the same state, time and parameters, and both continue identically
(navier/rbc2d_base.py:205-257, pypde/field.py:83-173,440-459)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GROUPS = ("temp", "ux", "uy", "pres")


def host(t):
    return t.detach().cpu().numpy()


def make(ra=1e5, **kw):
    from pypde_amd.navier import NavierStokes
    args = dict(case="rbc", shape=(20, 16), ra=ra, pr=1.0, dt=0.01, tsave=None, dealias=True,
                integrator="eu", beta=1.0, aspect=1.0)
    args.update(kw)
    return NavierStokes(**args)


def test_write_read_round_trip(tmp_path):
    from pypde_amd.h5io import H5File
    ns = make()
    ns.set_temperature(amplitude=0.2)
    for _ in range(30):
        ns.update()
        ns.update_time()
    fn = str(tmp_path / "state.h5")
    ns.write(filename=fn)
    with H5File(fn, "r") as hf:
        keys = set(hf.keys())
        assert keys == set(GROUPS) | {g + "/" + k for g in GROUPS for k in ("v", "vhat")} | \
            {"time", "x", "y", "nu", "kappa", "ra", "pr"}
        assert abs(float(hf.read("time")) - 0.3) < 1e-12
        assert abs(float(hf.read("ra")) - 1e5) < 1e-6
        # v is the physical field of vhat (MultiField.write transforms first)
        for g, f in zip(GROUPS, (ns.T, ns.U, ns.V, ns.pres)):
            assert np.array_equal(hf.read(g + "/vhat"), host(f.vhat))
            assert np.allclose(hf.read(g + "/v"), host(f.backward(f.vhat)), atol=1e-14)
        assert np.array_equal(hf.read("x"), np.asarray(ns.T.x))
    ns2 = make()
    ns2.read(filename=fn)
    assert ns2.fused
    assert abs(ns2.time - ns.time) < 1e-15
    for f in ("T", "U", "V", "pres"):
        assert np.array_equal(host(getattr(ns2, f).vhat), host(getattr(ns, f).vhat)), f
    for _ in range(10):
        ns.update()
        ns2.update()
    for f in ("T", "U", "V", "pres"):
        assert np.array_equal(host(getattr(ns2, f).vhat), host(getattr(ns, f).vhat)), f


def test_write_read_from_Ra_and_default_names(tmp_path, monkeypatch):
    ns = make(ra=2e4)
    ns.set_temperature(amplitude=0.1)
    ns.update()
    ns.write_from_Ra(folder=str(tmp_path))
    fn = os.path.join(str(tmp_path), "Flow_Ra2.000e+04.h5")
    assert os.path.exists(fn)
    ns2 = make(ra=2e4)
    ns2.read_from_Ra(folder=str(tmp_path))
    assert np.array_equal(host(ns2.T.vhat), host(ns.T.vhat))
    # filename=None: leading_str + _{time:07.2f}.h5 in the working directory
    monkeypatch.chdir(tmp_path)
    ns.T.t = ns.U.t = ns.V.t = ns.pres.t = 1.25
    ns.write(leading_str="flow")
    assert os.path.exists(os.path.join(str(tmp_path), "flow_0001.25.h5"))


def test_read_applies_file_parameters_in_reference_order(tmp_path):
    """read() recomputes nu/kappa before it takes ra/pr from the file
    (rbc2d_base.py:218-227): ra follows the file, nu the solver's previous ra."""
    from pypde_amd.navier.rbc2d_base import nu
    ns = make(ra=1e5)
    fn = str(tmp_path / "s.h5")
    ns.write(filename=fn)
    ns2 = make(ra=3e5)
    ns2.read(filename=fn)
    assert abs(ns2.ra - 1e5) < 1e-6
    assert ns2.nu == nu(3e5, 1.0, 1.0)
    assert ns2.engine.nu[0] == ns2.nu
