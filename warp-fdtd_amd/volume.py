"""Passive full-volume phasors at fixed physical points (drop-in for the
monitor part of warp_fdtd/volume.py:60-148).

All six components are sampled trilinearly at the same physical point; H is
centred in time with hprev before the window-weighted accumulation.  The
monitor reads hprev everywhere, so a solver carrying one saves the full
volume (Solver.hprev_box).
"""

import numpy as np

from ._lib import call, ptr, stream
from .array import Array
from .kernels import Kernel


def _record_volume(e, h, prev, positions, weights, step, phasor, nx, ny, nz, first, dim):
    call("wfd_record_volume", ptr(e), ptr(h), ptr(prev), ptr(positions), ptr(weights),
         ptr(step), ptr(phasor), int(dim), int(nx), int(ny), int(nz), int(first), stream())


# volume.py:274-302 record_volume; the point count is the launch dim.
record_volume = Kernel("wfd_record_volume", _record_volume, takes_dim=True)


class VolumeMonitor:
    """volume.py:91-148"""

    def __init__(self, solver, axes_m, window_monitor):
        self.axes_m = tuple(np.asarray(a, dtype=np.float64) for a in axes_m)
        if len(self.axes_m) != 3 or any(
                a.ndim != 1 or not a.size or not np.isfinite(a).all() or np.any(np.diff(a) <= 0)
                for a in self.axes_m):
            raise ValueError("Three finite increasing sampling axes required")
        self.shape = tuple(len(a) for a in self.axes_m)
        self.n = int(np.prod(self.shape))
        self.obj = {"class": "VolumeDiagnostic", "name": "volume"}
        self.name = "volume"
        self.first = window_monitor.first
        self.weights = window_monitor.weights
        self.exact = True
        dx = solver.meta["settings"]["spacing_nm"] * 1e-9
        axes = [(a - float(solver.data[c + "_edges_m"][0])) / dx for a, c in zip(self.axes_m, "xyz")]
        if any(np.any(a < 0) or np.any(a > n) for a, n in zip(axes, solver.shape)):
            raise ValueError("Sampling points must lie within the physical grid")
        points = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 3)
        self.positions = Array.from_numpy(points.astype("f"), np.float32, solver.device)
        self.phasor = Array.zeros((6, self.n, 2), device=solver.device)
        solver.monitors.append(self)
        solver.save_h_from = min(solver.save_h_from, self.first)

    def active_steps(self, steps):
        active = np.zeros(steps, bool)
        active[self.first:] = True
        return active

    def launch(self, solver):
        _record_volume(solver.E, solver.H, solver.hprev, self.positions, self.weights,
                       solver.record_step, self.phasor, *solver.shape, self.first, dim=self.n)

    def reset(self):
        self.phasor.zero_()

    def phasors(self):
        raw = self.phasor.numpy()
        return (raw[..., 0] + 1j * raw[..., 1]).reshape(6, *self.shape)
