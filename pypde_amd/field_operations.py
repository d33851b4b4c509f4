"""Field operations on the GPU (drop-in for pypde/field_operations.py:8-220)."""
import numpy as np
import torch

from . import ops
from .field import Field, FieldBC


def grad(field, deriv, return_field=False, scale=None):
    """Derivative of a field, per axis, divided by scale**deriv
    (field_operations.py:8-56).  Galerkin axes are mapped to Chebyshev
    coefficients even for deriv 0, as the reference does; the stencil, the
    recurrence and the rescale run in one line kernel per axis."""
    assert isinstance(field, (Field, FieldBC))
    if isinstance(deriv, int):
        deriv = (deriv,)
    assert field.ndim == len(deriv)
    dvhat = field.vhat
    for axis in range(field.ndim):
        sp = 1.0
        if scale is not None:
            assert len(scale) == field.ndim
            sp = scale[axis] ** deriv[axis]
        dvhat = field.derivative(dvhat, deriv[axis], axis=axis, scale_pow=sp)
    if return_field:
        xs = [b.family if hasattr(b, "family") else b for b in field.xs]
        out = Field(xs)
        out.vhat = dvhat
        return out
    return dvhat


def cheby_to_galerkin(uhat, galerkin_field):
    """field_operations.py:59-68."""
    for axis in range(uhat.dim()):
        b = galerkin_field.xs[axis]
        if hasattr(b, "from_chebyshev"):
            uhat = b.from_chebyshev(uhat, axis=axis)
    return uhat


def galerkin_to_cheby(vhat, galerkin_field):
    """field_operations.py:71-80."""
    for axis in range(vhat.dim()):
        b = galerkin_field.xs[axis]
        if hasattr(b, "to_chebyshev"):
            vhat = b.to_chebyshev(vhat, axis=axis)
    return vhat


def conv_term(v_field, u, deriv, deriv_field=None, dealias=False, scale=None):
    """u * d(v)/d(deriv) on the (dealiased) grid (field_operations.py:83-129)."""
    from .bases import Base
    assert isinstance(v_field, Field), "v_field must be instance Field"
    if deriv_field is None:
        if dealias:
            deriv_field = Field([Base(v_field.shape[0], "CH", dealias=3 / 2),
                                 Base(v_field.shape[1], "CH", dealias=3 / 2)])
        else:
            deriv_field = Field([Base(v_field.shape[0], "CH"), Base(v_field.shape[1], "CH")])
    vhat = grad(v_field, deriv, return_field=False, scale=scale)
    if dealias:
        dvdx = deriv_field.dealias.backward(vhat)
    else:
        dvdx = deriv_field.backward(vhat)
    return dvdx * u


def convective_term(v_field, ux, uz, deriv_field=None, add_bc=None, dealias=False, scale=None):
    """ux*dv/dx + uz*dv/dz (+add_bc), forward transformed (field_operations.py:132-169)."""
    conv = conv_term(v_field, ux, (1, 0), deriv_field, dealias, scale=scale)
    conv = conv + conv_term(v_field, uz, (0, 1), deriv_field, dealias, scale=scale)
    if add_bc is not None:
        conv = conv + add_bc
    if dealias:
        return deriv_field.dealias.forward(conv)
    return deriv_field.forward(conv)


def _w(d, like):
    return torch.as_tensor(np.asarray(d), dtype=ops.DTYPE, device=like.device)


def avg_x(f, dx):
    """field_operations.py:175-176."""
    w = _w(dx, f)
    return (f * w[:, None]).sum(dim=0) / w.sum()


def avg_vol(f, dx, dy):
    """field_operations.py:179-181."""
    wx, wy = _w(dx, f), _w(dy, f)
    favgx = (f * wx[:, None]).sum(dim=0) / wx.sum()
    return (favgx * wy).sum() / wy.sum()


def interpolate(Field_old, Field_new, spectral=True):
    """Spectral (or physical) zero-pad / truncate interpolation (field_operations.py:184-220)."""
    F_old = Field_old.vhat if spectral else Field_old.v
    F_new = Field_new.vhat if spectral else Field_new.v
    if F_old.dim() != F_new.dim():
        raise ValueError("Field must be of same dimension!")
    sl = tuple(slice(0, min(a, b)) for a, b in zip(F_old.shape, F_new.shape))
    F_new.zero_()
    F_new[sl] = F_old[sl]
    if spectral:
        Field_new.backward()
    else:
        Field_new.forward()
