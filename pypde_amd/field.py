"""Fields on the GPU: Field, FieldBC, MultiField.

Drop-in for pypde.field (pypde/field.py:9-465) and the per-axis driver of
pypde.bases.spectralspace (spectralspace.py:4-99).  `v` (physical values) and
`vhat` (spectral coefficients) are float64 device tensors.  Transforms along
axis 1 run directly on rows (no swapaxes copies); the dealiased sub-field pads
before backward and truncates after forward exactly as field.py:14-56 does,
with the padding folded into the transform matrices.
"""
import numpy as np
import torch

from . import ops
from .bases import MetaBase


class SpectralSpace:
    """N-D space from a list of bases (pypde/bases/spectralspace.py:4-66)."""

    def __init__(self, bases):
        if isinstance(bases, MetaBase):
            bases = [bases]
        assert all(isinstance(i, MetaBase) for i in bases)
        self._set_bases(bases)
        self.ndim = len(self.shape_physical)
        self.shape = self.shape_physical

    def _set_bases(self, bases):
        self.xs = list(bases)
        self.shape_physical = tuple(b.N for b in self.xs)
        self.shape_spectral = tuple(b.M for b in self.xs)

    def forward_fft(self, v, axis):
        assert isinstance(axis, int)
        return self.xs[axis].forward_fft(v, axis=axis)

    def backward_fft(self, vhat, axis):
        assert isinstance(axis, int)
        return self.xs[axis].backward_fft(vhat, axis=axis)

    def derivative(self, vhat, deriv, axis, out_cheby=True, scale_pow=1.0):
        return self.xs[axis].derivative(vhat, deriv, out_cheby, axis=axis, scale_pow=scale_pow)


class SpectralSpaceBC(SpectralSpace):
    """Space of the inhomogeneous part (spectralspace.py:69-99): the BC basis
    along `axis`, the Chebyshev family along the others."""

    def __init__(self, bases, axis):
        SpectralSpace.__init__(self, bases)
        self.axis = axis
        for i in range(self.ndim):
            if i == axis:
                if hasattr(self.xs[i], "bc"):
                    self.xs[i] = self.xs[i].bc
            elif hasattr(self.xs[i], "family"):
                self.xs[i] = self.xs[i].family
        SpectralSpace.__init__(self, self.xs)


def _dx(x):
    xm = np.zeros(x.size + 1)
    xm[0], xm[-1] = x[0], x[-1]
    xm[1:-1] = (x[1:] + x[:-1]) / 2.0
    return np.diff(xm)


class FieldBase:
    """forward/backward shared by Field and FieldBC (field.py:9-80)."""

    def forward(self, v=None, undealias_after=None):
        vhat = self.v if v is None else ops.to_dev(v)
        if undealias_after is None:
            undealias_after = self.dealiased_space
        for axis in range(self.ndim):
            b = self.xs[axis]
            if undealias_after and not hasattr(b, "family"):
                # Chebyshev: truncation folded into the forward matrix
                vhat = b.forward_fft(vhat, axis=axis, n_coef=self.size_undealiased[axis])
            else:
                vhat = self.forward_fft(vhat, axis=axis)
                if undealias_after:
                    vhat = _unpad(vhat, self.size_undealiased[axis], axis)
        if v is None:
            self.vhat = vhat
        else:
            return vhat

    def backward(self, vhat=None, dealias_before=None):
        v = self.vhat if vhat is None else ops.to_dev(vhat)
        # zero padding to xs[axis].M is implicit: the backward matrices take
        # any number of leading coefficients
        for axis in range(self.ndim):
            v = self.backward_fft(v, axis=axis)
        if vhat is None:
            self.v = v
        else:
            return v

    @property
    def x(self):
        return self.xs[0].x

    @property
    def y(self):
        if self.ndim < 2:
            raise ValueError("Dimension y not defined for ndim<2.")
        return self.xs[1].x

    @property
    def dx(self):
        return _dx(self.x)

    @property
    def dy(self):
        return _dx(self.y)

    # -- read / write (field.py:83-173): HDF5 through pypde_amd.h5io
    @staticmethod
    def _filename(filename, leading_str, add_time, t):
        if filename is None:
            filename = leading_str
            if add_time:
                filename = filename + "_{:07.2f}".format(float(t))
            filename = filename + ".h5"
        return filename

    def write(self, filename="file_0.h5", dict=None, leading_str="flow", add_time=True,
              grp_name=""):
        """Datasets {grp}/v, {grp}/vhat, time, x, y and the entries of dict
        (created, or overwritten in place when they exist)."""
        from .h5io import H5File
        filename = self._filename(filename, leading_str, add_time, self.t)
        if grp_name and grp_name[-1] != "/":
            grp_name = grp_name + "/"
        print("Write {:s} ...".format(filename))
        with H5File(filename, "a") as hf:
            self.write_single_hdf5(hf, grp_name + "v", self.v)
            self.write_single_hdf5(hf, grp_name + "vhat", self.vhat)
            self.write_single_hdf5(hf, "time", self.t)
            if self.x is not None:
                self.write_single_hdf5(hf, "x", self.x)
            if self.y is not None:
                self.write_single_hdf5(hf, "y", self.y)
            if dict is not None:
                for key in dict:
                    self.write_single_hdf5(hf, key, dict[key])

    @staticmethod
    def write_single_hdf5(hf, name, data):
        hf.write(name, data)

    @staticmethod
    def read_single_hdf5(hf, name):
        data = hf.read(name)
        if data is not None:
            return data
        print("Cannot read {:} ...".format(name))
        print("Following keys exist: ")
        for key in hf.keys():
            print_grp(key)
        return 0.0

    def read(self, filename="file_0.h5", dict=None, leading_str="flow", add_time=True,
             grp_name=""):
        """Fill v, vhat (in place, so views bound elsewhere see the values),
        t and the entries of dict from the file; a missing dataset reads 0.0."""
        from .h5io import H5File
        filename = self._filename(filename, leading_str, add_time, self.t)
        print("Read {:s} ...".format(filename))
        if grp_name and grp_name[-1] != "/":
            grp_name = grp_name + "/"
        with H5File(filename, "r") as hf:
            for attr in ("v", "vhat"):
                dst = getattr(self, attr)
                val = self.read_single_hdf5(hf, grp_name + attr)
                if np.ndim(val) == 0:
                    dst.fill_(float(val))
                else:
                    if tuple(val.shape) != tuple(dst.shape):
                        raise ValueError("could not broadcast %s of shape %s into %s"
                                         % (grp_name + attr, val.shape, tuple(dst.shape)))
                    dst.copy_(torch.as_tensor(val))
            t = self.read_single_hdf5(hf, "time")
            self.t = float(t)
            if dict is not None:
                for key in dict:
                    dict[key] = self.read_single_hdf5(hf, key)


def print_grp(name):
    print("Group:", name)


def _unpad(a, n, axis):
    if a.dim() == 1:
        return a[:n].contiguous()
    return (a[:n] if axis == 0 else a[:, :n]).contiguous()


class Field(SpectralSpace, FieldBase):
    """Physical / spectral field on a product of bases (field.py:180-356)."""

    def __init__(self, bases):
        if isinstance(bases, MetaBase):
            bases = [bases]
        SpectralSpace.__init__(self, bases)
        self.bases = bases
        dev = ops.device()
        self.v = torch.zeros(self.shape_physical, dtype=ops.DTYPE, device=dev)
        self.vhat = torch.zeros(self.shape_spectral, dtype=ops.DTYPE, device=dev)
        self.field_bc = None
        self.t = 0
        self.V = []
        self.T = []
        self.dealiased_space = False
        self.create_dealiased_field(bases)

    def create_dealiased_field(self, bases):
        if all(hasattr(i, "dealias") for i in bases):
            self.dealias = Field([i.dealias for i in bases])
            self.dealias.size_undealiased = [self.xs[i].M for i in range(self.ndim)]
            self.dealias.dealiased_space = True

    def add_field_bc(self, field_bc):
        assert isinstance(field_bc, FieldBC)
        self.field_bc = field_bc

    def make_homogeneous(self):
        import warnings
        if self.field_bc is None:
            warnings.warn("No inhomogeneous field found. Call add_bc(bchat,axis) first!")
        else:
            assert self.v.shape == self.inhomogeneous.shape, "Shape mismatch in make_homogeneous"
        return self.v - self.inhomogeneous

    @property
    def total(self):
        return self.homogeneous + self.inhomogeneous

    @property
    def homogeneous(self):
        return self.v

    @property
    def inhomogeneous(self):
        if self.field_bc is not None:
            return self.field_bc.v
        return 0

    def save(self, transform=True):
        if transform:
            self.backward()
        self.V.append(self.v.detach().cpu().numpy().copy())
        self.T.append(self.t)


class FieldBC(SpectralSpaceBC, FieldBase):
    """Inhomogeneous part from boundary values (field.py:359-414)."""

    def __init__(self, bases, axis):
        SpectralSpaceBC.__init__(self, bases, axis)
        dev = ops.device()
        self.v = torch.zeros(self.shape_physical, dtype=ops.DTYPE, device=dev)
        self.vhat = torch.zeros(self.shape_spectral, dtype=ops.DTYPE, device=dev)
        self.dealiased_space = False

    def add_bc(self, bc):
        expected = list(self.shape_physical)
        expected[self.axis] = self.shape_spectral[self.axis]
        bc = ops.to_dev(bc)
        assert tuple(bc.shape) == tuple(expected)
        self.v = self.backward_fft(bc, axis=self.axis)
        self.forward()


class MultiField:
    """Collection of fields with collective save / time update (field.py:417-465)."""

    def __init__(self, fields, names):
        self.fields, self.names = [], []
        for f, n in zip(fields, names):
            if not isinstance(f, Field):
                raise ValueError("Must be of type Field.")
            self.fields.append(f)
            self.names.append(n)

    def save(self):
        for f in self.fields:
            f.save()

    def update_time(self, dt):
        for f in self.fields:
            f.t += dt

    def read(self, filename=None, leading_str="", add_time=True, dict={}):
        for f, n in zip(self.fields, self.names):
            f.read(filename=filename, leading_str=leading_str, add_time=add_time, dict=dict,
                   grp_name=n)

    def write(self, filename=None, leading_str="", add_time=True, dict={}):
        for f, n in zip(self.fields, self.names):
            f.backward()
            f.write(filename=filename, leading_str=leading_str, add_time=add_time, dict=dict,
                    grp_name=n)
