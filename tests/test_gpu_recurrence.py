"""The kernel forms of pypde_recurrence_solve (block form with 64 lines per
block, variant 1; lane-chunk form, variant 2; block form with 32 lines per
block, variant 3, and with 32 positions per chunk, variant 4) against a sequential evaluation of the normalised
recurrence, for every table combination the plans use, both axes, line lengths
around the chunk sizes of either form, per-entry and per-line tables (both
layouts), in place and out of place."""
import ctypes

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

FORMS = {
    "fdma": ("LM", "R", "MU", "NU"),          # Helmholtz four-diagonal sweeps
    "proj": ("pre", "F", "LM", "MU"),          # Galerkin projection (S^T S)^-1 S^T
    "twodma": ("R", "MU"),                     # upper two-band solve
    "full": ("F", "LM", "R", "MU", "NU"),
}


def tables(rng, n, names, lead=()):
    """Random tables with the documented zero padding of chain starts / ends;
    multipliers small enough that both chains stay well conditioned."""
    t = {}
    for k in names:
        if k in ("F", "R"):
            a = 1.0 + 0.3 * rng.standard_normal(lead + (n,))
        else:
            a = 0.45 * rng.uniform(-1, 1, lead + (n,))
        if k == "LM":
            a[..., :2] = 0.0
        if k == "MU":
            a[..., max(n - 2, 0):] = 0.0
        if k == "NU":
            a[..., max(n - 4, 0):] = 0.0
        t[k] = a
    return t


def reference(x, t, n):
    """Sequential evaluation along axis 0 of x [n_in, m]; tables [n] or [n, m]."""
    def col(k):
        a = t[k]
        return a[:, None] if a.ndim == 1 else a
    y = x[:n].copy()
    if "pre" in t:
        y = y + col("pre") * x[2:n + 2]
    if "LM" in t:
        F = col("F") if "F" in t else np.ones((n, 1))
        LM = col("LM")
        for i in range(n):
            y[i] = F[i] * y[i] + (LM[i] * y[i - 2] if i >= 2 else 0.0)
    R = col("R") if "R" in t else np.ones((n, 1))
    MU = col("MU")
    NU = col("NU") if "NU" in t else None
    z = np.zeros_like(y)
    for i in range(n - 1, -1, -1):
        z[i] = R[i] * y[i]
        if i + 2 < n:
            z[i] = z[i] + MU[i] * z[i + 2]
        if NU is not None and i + 4 < n:
            z[i] = z[i] + NU[i] * z[i + 4]
    return z


def run(ops, lib, Recurrence, x, out, axis, n_in, n, dev, variant, per_line, entry_stride=0):
    g = ops._geom(ops._as3(x), ops._as3(out), axis, n_in, n)
    r = Recurrence()
    for k, v in dev.items():
        setattr(r, k, v.data_ptr())
    r.entry_stride, r.per_line, r.variant = entry_stride, per_line, variant
    ops.check(lib.pypde_recurrence_solve(ctypes.byref(g), ctypes.byref(r), ops.stream_ptr()),
              "pypde_recurrence_solve")


@pytest.mark.parametrize("form", sorted(FORMS))
@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("variant", [1, 2, 3, 4])
def test_shared_and_per_entry_tables(form, axis, variant):
    import torch
    from pypde_amd import ops
    from pypde_amd._lib import Recurrence, lib
    names = FORMS[form]
    rng = np.random.default_rng(sorted(FORMS).index(form) * 8 + axis * 4 + variant)
    for n in (3, 5, 64, 127, 128, 129, 200, 256, 257, 400, 512):
        pre = "pre" in names
        n_in = n + 2 if pre else n
        nl, B = 37, 3
        # per-entry tables (entry_stride = one table set)
        t = tables(rng, n, names, lead=(B,))
        stacked = np.zeros((B, len(names), n))
        for j, k in enumerate(names):
            stacked[:, j] = t[k]
        st = ops.to_dev(stacked)
        dev = {k: st[0, j] for j, k in enumerate(names)}
        x = rng.standard_normal((B, n_in, nl) if axis == 0 else (B, nl, n_in))
        xd = ops.to_dev(x)
        out = torch.zeros((B, n, nl) if axis == 0 else (B, nl, n), dtype=torch.float64,
                          device=xd.device)
        run(ops, lib, Recurrence, xd, out, axis, n_in, n, dev, variant, 0,
            entry_stride=len(names) * n)
        got = out.cpu().numpy()
        for e in range(B):
            te = {k: t[k][e] for k in names}
            xe = x[e] if axis == 0 else x[e].T
            ref = reference(xe, te, n)
            ge = got[e] if axis == 0 else got[e].T
            assert rel(ge, ref) < 1e-13, (form, axis, variant, n, e)
        # shared tables, in place where the form allows
        if not pre:
            t1 = {k: t[k][0] for k in names}
            dev1 = {k: ops.to_dev(t1[k]) for k in names}
            xd = ops.to_dev(x)
            run(ops, lib, Recurrence, xd, xd, axis, n, n, dev1, variant, 0)
            got = xd.cpu().numpy()
            for e in range(B):
                xe = x[e] if axis == 0 else x[e].T
                ge = got[e] if axis == 0 else got[e].T
                assert rel(ge, reference(xe, t1, n)) < 1e-13, (form, axis, variant, n, "inplace")


@pytest.mark.parametrize("variant", [1, 2, 3, 4])
def test_per_line_tables(variant):
    """The Poisson form with tables per line: [n][nlines] for the block forms,
    line-major [nlines][n] for the lane-chunk form (axis 0, lines = columns)."""
    import torch
    from pypde_amd import ops
    from pypde_amd._lib import Recurrence, lib
    rng = np.random.default_rng(7 + variant)
    names = FORMS["fdma"]
    for n, nl in ((3, 5), (64, 64), (129, 70), (254, 254), (400, 33), (512, 9)):
        t = tables(rng, n, names, lead=(nl,))          # [nlines][n]
        if variant != 2:
            block = ops.to_dev(np.stack([t[k].T for k in names]))   # [4][n][nlines]
            layout = 1
        else:
            block = ops.to_dev(np.stack([t[k] for k in names]))     # [4][nlines][n]
            layout = 2
        dev = {k: block[j] for j, k in enumerate(names)}
        x = rng.standard_normal((2, n, nl))
        xd = ops.to_dev(x)
        run(ops, lib, Recurrence, xd, xd, 0, n, n, dev, variant, layout)
        got = xd.cpu().numpy()
        tl = {k: t[k].T for k in names}                 # [n][nlines] for the reference
        for e in range(2):
            assert rel(got[e], reference(x[e], tl, n)) < 1e-13, (variant, n, nl, e)


def test_layout_and_variant_checks():
    """Mismatched layout / variant pairs are rejected before any launch."""
    from pypde_amd import ops
    from pypde_amd._lib import PypdeError, Recurrence, lib
    import torch
    x = torch.zeros((8, 4), dtype=torch.float64, device="cuda")
    t = torch.zeros((4, 8, 4), dtype=torch.float64, device="cuda")
    dev = {k: t[j] for j, k in enumerate(FORMS["fdma"])}
    for per_line, variant in ((1, 2), (3, 0), (0, 5), (2, 3), (2, 4), (0, -1)):
        with pytest.raises(PypdeError):
            run(ops, lib, Recurrence, x, x, 0, 8, 8, dev, variant, per_line)
