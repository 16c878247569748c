"""Yee half-step kernels with the argument order of warp_fdtd/kernels.py.

``update_e``/``update_h`` (and the ``_packed`` variants) take the Warp kernel's
inputs (f, g, inv_eps, psi, coeff, source, waves, step, hprev, save_h_from,
nx, ny, nz, source_x, courant, periodic) and enqueue the HIP kernel of the
same name; ``launch(kernel, dim, inputs=[...])`` mirrors ``wp.launch``.
"""

import ctypes

import numpy as np

from . import _lib
from ._lib import call, ptr, stream


class Kernel:
    """A libwfdtd entry point taking the reference kernel's inputs."""

    def __init__(self, name, fn, dim=None, takes_dim=False):
        self.name = name
        self._fn = fn
        self._dim = dim
        self.takes_dim = takes_dim

    def __call__(self, *inputs, **kw):
        self._fn(*inputs, **kw)

    def expected_dim(self, inputs):
        return None if self._dim is None else self._dim(inputs)

    def __repr__(self):
        return f"<libwfdtd kernel {self.name}>"


def launch(kernel, dim, inputs, device=None):
    """wp.launch(kernel, dim, inputs=...) for libwfdtd kernels.  Kernels whose
    extent is only given by the launch (plane detectors, adjoint box kernels)
    receive it as ``dim``; the others check it against their grid."""
    size = dim if isinstance(dim, int) else int(np.prod(dim))
    if kernel.takes_dim:
        kernel(*inputs, dim=size)
        return
    expected = kernel.expected_dim(inputs)
    if expected is not None and size != expected:
        raise ValueError(f"{kernel.name}: launch dim {dim} != {expected}")
    kernel(*inputs)


def _rows(coeff):
    return coeff.shape[1]


def _stride(a):
    return a.shape[1] if len(a.shape) > 1 else 1


def _make_update(name):
    def fn(f, g, inv_eps, psi, coeff, source, waves, step, hprev, save_h_from,
           nx, ny, nz, source_x, courant, periodic):
        call(name, ptr(f), ptr(g), ptr(inv_eps), ptr(psi), _stride(psi), ptr(coeff),
             _rows(coeff), ptr(source), _stride(source), ptr(waves), _stride(waves),
             ptr(step), ptr(hprev), int(save_h_from), int(nx), int(ny), int(nz),
             int(source_x), float(courant), int(periodic), stream())

    return Kernel(name, fn, dim=lambda inputs: int(inputs[10]) * int(inputs[11]) * int(inputs[12]))


update_e = _make_update("wfd_update_e")
update_h = _make_update("wfd_update_h")
update_e_packed = _make_update("wfd_update_e_packed")
update_h_packed = _make_update("wfd_update_h_packed")


def _advance(step):
    call("wfd_advance", ptr(step), 1, stream())


advance = Kernel("wfd_advance", _advance, dim=lambda inputs: 1)


def yee_step(plan):
    """One fused E+H step described by a wfd_yee_plan."""
    call("wfd_yee_step", ctypes.byref(plan), stream())


def synchronize():
    call("wfd_device_synchronize")


__all__ = ["update_e", "update_h", "update_e_packed", "update_h_packed", "advance",
           "launch", "yee_step", "synchronize", "_lib"]
