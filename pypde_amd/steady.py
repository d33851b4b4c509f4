"""Steady states of the time-step map by Jacobian-free Newton-Krylov on the GPU.

The reference finds steady states with scipy's root(method="krylov") on
steady_fun (navier/rbc2d_base.py:259-341): F(X) = (step(X) - X)/dt for
X = [T, U, V].  Its F also advances the pressure held in the solver object, so
F depends on the call history as well as on X, and the Newton iteration stalls
(20x16, Ra=1e4: |F| stays at 1.5 after 100 iterations).  Here the pressure is
part of the unknown, z = [T, U, V, pres], and F(z) = (Phi(z) - z)/tau with Phi
n steps of the fused engine from exactly z (tau = n dt).  The fixed points of
that map are the attractors that time integration converges to.

Everything stays on the device: the state vectors, the Krylov basis
(a [m+1, n] tensor) and the Gram-Schmidt projections (two block passes per
Arnoldi step).  Only the (m+1)-long Hessenberg columns go to the host, where
the Givens rotations run.
"""
import math

import numpy as np
import torch


class SteadyResult(dict):
    """Attribute access to the solver's output, like scipy's OptimizeResult:
    x, success, status, message, nit (Newton iterations), nfev (F evaluations),
    fun (F at x), fnorm (max |F| at x), history (max |F| per iteration)."""

    def __getattr__(self, name):
        try:
            return self[name]
        except KeyError as exc:
            raise AttributeError(name) from exc

    def __setattr__(self, name, value):
        self[name] = value


def gmres(matvec, b, tol=1e-6, restart=30, maxiter=20, x0=None):
    """Restarted GMRES(restart) for A x = b with A given by matvec (device
    tensors).  Stops when |b - A x|_2 <= tol |b|_2 by the Arnoldi residual
    estimate, or after maxiter restart cycles.  Returns (x, iterations, rel. residual)."""
    bnorm = float(torch.linalg.vector_norm(b))
    x = torch.zeros_like(b) if x0 is None else x0.clone()
    if bnorm == 0.0:
        return x, 0, 0.0
    n = b.numel()
    V = torch.empty((restart + 1, n), dtype=b.dtype, device=b.device)
    its, rel = 0, 1.0
    for _ in range(maxiter):
        r = b - matvec(x) if (x0 is not None or its > 0) else b.clone()
        beta = float(torch.linalg.vector_norm(r))
        rel = beta / bnorm
        if rel <= tol:
            break
        V[0] = r / beta
        H = np.zeros((restart + 1, restart))
        cs, sn = np.zeros(restart), np.zeros(restart)
        g = np.zeros(restart + 1)
        g[0] = beta
        k = 0
        for j in range(restart):
            w = matvec(V[j])
            Vj = V[: j + 1]
            h = Vj @ w
            w = w - Vj.t() @ h
            h2 = Vj @ w  # second pass: orthogonality to roundoff
            w = w - Vj.t() @ h2
            hn = torch.cat([h + h2, torch.linalg.vector_norm(w).reshape(1)]).cpu().numpy()
            H[: j + 2, j] = hn
            if hn[j + 1] > 0:
                V[j + 1] = w / hn[j + 1]
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            d = math.hypot(H[j, j], H[j + 1, j])
            cs[j], sn[j] = (1.0, 0.0) if d == 0 else (H[j, j] / d, H[j + 1, j] / d)
            H[j, j], H[j + 1, j] = d, 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            k = j + 1
            its += 1
            rel = abs(g[j + 1]) / bnorm
            if rel <= tol or hn[j + 1] == 0:
                break
        # least squares on the rotated triangle: well defined when J is singular
        y = np.linalg.lstsq(np.triu(H[:k, :k]), g[:k], rcond=None)[0]
        x = x + V[:k].t() @ torch.as_tensor(y, dtype=b.dtype, device=b.device)
        if rel <= tol:
            break
    return x, its, rel


def newton_krylov(F, z0, fatol=1e-8, maxiter=300, inner_m=100, inner_maxiter=30,
                  rdiff=None, eta_max=0.9, disp=False):
    """Inexact Newton for F(z) = 0 (z, F(z) device tensors).

    Steps solve J d = -F(z) by GMRES(inner_m) with at most inner_maxiter
    restart cycles, J v = (F(z + e v) - F(z))/e with e = rdiff max(1, |z|)/|v|
    (rdiff = sqrt of the unit roundoff by default), to the Eisenstat-Walker
    relative tolerance eta_k = 0.9 (|F_k| / |F_k-1|)^2, safeguarded and capped at
    eta_max.  Each step is backtracked (Armijo, quadratic interpolation)
    on |F|_2.  Converged when max|F| <= fatol, the norm scipy's fatol uses."""
    eps = float(torch.finfo(z0.dtype).eps)
    rdiff = math.sqrt(eps) if rdiff is None else rdiff
    z = z0.clone()
    Fz = F(z)
    nfev = 1
    fmax = float(Fz.abs().max())
    fn = float(torch.linalg.vector_norm(Fz))
    hist = [fmax]
    eta, fn_old = 0.5, None
    status, message, nit = 1, "The maximum number of iterations allowed has been reached.", 0
    if not math.isfinite(fmax):
        status, message = 3, "F is not finite at the starting point."
    for it in range(maxiter):
        if fmax <= fatol or status == 3:
            if status != 3:
                status, message = 0, "A solution was found at the specified tolerance."
            break
        if fn_old is not None:
            e_new = 0.9 * (fn / fn_old) ** 2
            if 0.9 * eta * eta > 0.1:
                e_new = max(e_new, 0.9 * eta * eta)
            eta = min(eta_max, e_new)
        eta = max(eta, 0.5 * fatol / max(fmax, 1e-300))
        znorm = float(torch.linalg.vector_norm(z))
        count = [0]

        def jv(v, z=z, Fz=Fz, znorm=znorm):
            vn = float(torch.linalg.vector_norm(v))
            if vn == 0.0:
                return torch.zeros_like(v)
            e = rdiff * max(1.0, znorm) / vn
            count[0] += 1
            return (F(z + e * v) - Fz) / e

        d, _, _ = gmres(jv, -Fz, tol=eta, restart=inner_m, maxiter=inner_maxiter)
        nfev += count[0]
        lam, phi0 = 1.0, fn * fn
        for _ in range(20):
            zt = z + lam * d
            Ft = F(zt)
            nfev += 1
            ft = float(torch.linalg.vector_norm(Ft))
            if math.isfinite(ft) and ft * ft <= (1.0 - 2e-4 * lam) * phi0:
                break
            a = (ft * ft - phi0 + 2.0 * lam * phi0) / (lam * lam) if math.isfinite(ft) else 0.0
            lam_q = phi0 / a if a > 0 else 0.1 * lam
            lam = min(0.5 * lam, max(0.1 * lam, lam_q))
        else:
            status, message = 2, "The line search failed to reduce |F|."
            nit = it + 1
            break
        fn_old, z, Fz, fn = fn, zt, Ft, ft
        fmax = float(Fz.abs().max())
        hist.append(fmax)
        nit = it + 1
        if disp:
            print("%d:  |F(x)| = %g; step %g" % (nit, fmax, lam))
    else:
        if fmax <= fatol:
            status, message = 0, "A solution was found at the specified tolerance."
    return SteadyResult(x=z, success=status == 0, status=status, message=message, nit=nit,
                        nfev=nfev, fun=Fz, fnorm=fmax, history=np.array(hist))
