"""Write tests/golden/rbc2d_c3_dense_oracle.npz: the c3 trajectory (512x512,
rbc, Ra=1e8, Pr=1, dt=1e-3, Euler, seed 0, noise 1e-3, 2000 steps) computed by
the oracle with its Chebyshev transforms evaluated as dense products with the
transform matrices instead of scipy's FFT-based DCT-I.  Both are exact
restatements of chebyshev.py:67-93 in FP64 that differ only in rounding, so the
distance between this series and the reference's own (rbc2d_c3_512.npz) is the
rounding floor of the trajectory at this size, against which the GPU series is
judged.

    python tests/golden/make_c3_floor.py        (about 30 min on 8 cores)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from oracle import pypde_oracle as o  # noqa: E402

_mats = {}


def _matrix(fn, basis, n):
    """The transform as a matrix: fn applied to the identity (column k = the
    image of the k-th unit vector)."""
    key = (fn.__name__, basis.N, n)
    if key not in _mats:
        _mats[key] = fn(basis, np.eye(n))
    return _mats[key]


_fwd, _bwd = o.Basis._cheb_forward, o.Basis._cheb_backward


def dense_forward(self, f):
    return _matrix(_fwd, self, f.shape[0]) @ f


def dense_backward(self, c):
    return _matrix(_bwd, self, c.shape[0]) @ c


def main(nsteps=2000):
    o.Basis._cheb_forward = dense_forward
    o.Basis._cheb_backward = dense_backward
    N = 512
    ns = o.NavierStokes(case="rbc", shape=(N, N), ra=1e8, pr=1.0, dt=1e-3, tsave=None)
    o.seeded_initial_state(ns, seed=0)
    nu, nv, ke = np.empty(nsteps), np.empty(nsteps), np.empty(nsteps)
    t0 = time.perf_counter()
    for n in range(nsteps):
        ns.update()
        nu[n], nv[n] = ns.eval_Nu()
        ke[n] = ns.eval_KE()
        if n % 100 == 0:
            print("step %d  %.1f s  Nu %.15f" % (n, time.perf_counter() - t0, nu[n]), flush=True)
    path = os.path.join(HERE, "rbc2d_c3_dense_oracle.npz")
    np.savez_compressed(path, nu_plate=nu, nu_vol=nv, ke=ke)
    print("wrote", path)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 2000)
