"""Colocated field/flux detectors on libwfdtd (drop-in for warp_fdtd/detectors.py).

Time accumulation runs on the device.  On the fused path the step kernel
stages each detector's raw footprint every step and one replay launch per
graph chunk records all detectors in step order (bitwise the per-step
result); otherwise every Monitor is recorded by one launch per step
(``record_all``).  ``record`` is the single-monitor kernel with the
reference's argument order.
"""

import ctypes

import numpy as np

from ._lib import MonitorDesc, ReplayDesc, WFD_MAX_MONITORS, call, ptr, stream
from .array import Array
from .kernels import Kernel


def _record(e, h, prev, step, weights, indices, phasor, flux, area, nx, ny, nz, x0, y0,
            z0, sy, sz, axis, exact, is_phasor, dim):
    if dim % (sy * sz):
        raise ValueError("record: launch dim must be a whole number of (sy, sz) planes")
    steps = min(len(indices), len(weights))
    call("wfd_record", ptr(e), ptr(h), ptr(prev), ptr(step), ptr(weights), ptr(indices),
         ptr(phasor), ptr(flux), ptr(area), int(nx), int(ny), int(nz), int(x0), int(y0),
         int(z0), int(dim // (sy * sz)), int(sy), int(sz), int(axis), int(exact),
         int(is_phasor), int(steps), stream())


# detectors.py:65-106 record; the plane's x extent comes from the launch dim.
record = Kernel("wfd_record", _record, takes_dim=True)


class Monitor:
    """detectors.py:109-211."""

    def __init__(self, solver, obj):
        self.obj = obj
        self.name = obj["name"]
        self.is_phasor = obj["class"] in ("ModeOverlapDetector", "PhasorDetector")
        if obj["class"] not in ("ModeOverlapDetector", "PhasorDetector",
                                "PoyntingFluxDetector"):
            raise ValueError(f"Unsupported detector {obj['class']}")
        self.shape = tuple(hi - lo for lo, hi in obj["grid_slice"])
        self.n = int(np.prod(self.shape))
        self.exact = self.is_phasor or solver.meta["settings"].get("exact_interpolation", False)
        indices = np.asarray(solver.data[self.name + "/_time_step_to_arr_idx"], np.int32)
        self.indices_numpy = indices
        self.indices = Array.from_numpy(indices, np.int32, solver.device)
        self.first = int(np.flatnonzero(indices >= 0)[0])
        weights = np.zeros(solver.steps, np.complex64)
        if self.is_phasor:
            src = next(o for o in solver.meta["objects"] if o["class"] == "ModePlaneSource")
            t = np.arange(solver.steps, dtype=np.float32) * np.float32(solver.meta["dt_s"])
            phase = np.float32(2 * np.pi * src["frequency_hz"]) * t
            weights = (np.exp(1j * phase).astype(np.complex64)
                       * np.float32(2 / obj["window_sum"])
                       * solver.data[self.name + "/_window_at_time_step_arr"])
        self.weights = Array.from_numpy(np.column_stack((weights.real, weights.imag)),
                                        np.float32, solver.device)
        self.phasor = Array.zeros((6, self.n if self.is_phasor else 1, 2), device=solver.device)
        self.flux = Array.zeros(int(indices.max()) + 1, device=solver.device)
        self.axis = obj.get("propagation_axis", 2)
        widths = []
        for a, axis in enumerate("xyz"):
            lo, hi = obj["grid_slice"][a]
            v = (np.diff(solver.data[axis + "_edges_m"])[lo:hi] if a != self.axis
                 else np.ones(hi - lo, np.float32))
            sh = [1, 1, 1]
            sh[a] = len(v)
            widths.append(v.reshape(sh))
        self.area_numpy = np.broadcast_to(widths[0] * widths[1] * widths[2], self.shape).copy()
        if obj["class"] == "ModeOverlapDetector":
            self.area_numpy /= self.area_numpy.mean()
        self.area = Array.from_numpy(self.area_numpy.ravel(), np.float32, solver.device)
        self.shape_of_grid = solver.shape

    def active_steps(self, steps):
        return self.indices_numpy[:steps] >= 0

    def footprint(self):
        return colocation_footprint(self.obj["grid_slice"], self.shape_of_grid)

    def stage_box(self):
        """Cells whose raw fields the step stages for this monitor."""
        if self.exact:
            return self.footprint()
        (x0, x1), (y0, y1), (z0, z1) = self.obj["grid_slice"]
        return (x0, x1, y0, y1, z0, z1)

    def replay_desc(self):
        d = ReplayDesc()
        d.stage = ptr(self.stage)
        for a, v in enumerate(self.stage_box()):
            d.box[a] = int(v)
        d.x0, d.y0, d.z0 = [a for a, b in self.obj["grid_slice"]]
        d.sx, d.sy, d.sz = self.shape
        d.kind = 0 if self.is_phasor else 1
        d.exact, d.axis = int(self.exact), int(self.axis)
        d.steps = int(len(self.indices_numpy))
        d.indices = ptr(self.indices)
        d.weights = ptr(self.weights)
        d.phasor = ptr(self.phasor)
        d.flux = ptr(self.flux)
        d.area = ptr(self.area)
        return d

    def reset(self):
        self.phasor.zero_()
        self.flux.zero_()

    def descriptor(self):
        lo = [a for a, b in self.obj["grid_slice"]]
        d = MonitorDesc()
        d.indices, d.weights = ptr(self.indices), ptr(self.weights)
        d.phasor, d.flux, d.area = ptr(self.phasor), ptr(self.flux), ptr(self.area)
        d.x0, d.y0, d.z0 = lo
        d.sx, d.sy, d.sz = self.shape
        d.axis, d.exact, d.is_phasor = int(self.axis), int(self.exact), int(self.is_phasor)
        d.steps = int(len(self.indices_numpy))
        return d

    def launch(self, s):
        record_all(s, [self])

    def phasors(self):
        raw = self.phasor.numpy()
        return (raw[..., 0] + 1j * raw[..., 1]).reshape(6, *self.shape)

    def power(self, s):
        ph = self.phasors()
        em = s.data[self.name + "/_mode_E"][0]
        hm = s.data[self.name + "/_mode_H"][0]
        integrand = (np.cross(em, ph[3:].conj(), axis=0)
                     + np.cross(ph[:3].conj(), hm, axis=0))[self.axis]
        alpha = np.sum(integrand * self.area_numpy) / 4
        return float(abs(alpha) ** 2)


def colocation_footprint(grid_slice, shape):
    """Cells of hprev the exact colocation of a plane reads: x-1..x, y-1..y,
    z..z+1 around every plane cell (detectors.py:49-56), clipped to the grid."""
    (x0, x1), (y0, y1), (z0, z1) = grid_slice
    return (max(x0 - 1, 0), min(x1, shape[0]), max(y0 - 1, 0), min(y1, shape[1]),
            max(z0, 0), min(z1 + 1, shape[2]))


C0 = 299792458.0


class DFTMonitor:
    """Running DFT of the colocated E and H on a plane at several wavelengths,
    accumulated in float64 (no reference counterpart; one wavelength with the
    reference's weights is the PhasorDetector of detectors.py:99-103).

    The fused step stages the plane's raw footprint every step and one replay
    per graph chunk colocates and folds the staged steps into the accumulators:
    phasor[f, c, i] += sum_t w[t, f] * field_c(t, i) with
    w[t, f] = exp(i 2 pi f t dt) dt.
    """

    is_phasor = True

    def __init__(self, solver, obj):
        self.obj = obj
        self.name = obj["name"]
        self.shape = tuple(hi - lo for lo, hi in obj["grid_slice"])
        self.lo = tuple(lo for lo, hi in obj["grid_slice"])
        self.n = int(np.prod(self.shape))
        self.axis = int(obj.get("propagation_axis", 0))
        self.exact = bool(obj.get("exact_interpolation", True))
        self.wavelengths = np.asarray(obj["wavelengths_m"], np.float64)
        self.frequencies = C0 / self.wavelengths
        self.nf = len(self.wavelengths)
        self.steps = solver.steps
        self.first = 0
        self.dt = float(solver.meta["dt_s"])
        t = np.arange(solver.steps, dtype=np.float64) * self.dt
        w = np.exp(2j * np.pi * self.frequencies[None, :] * t[:, None]) * self.dt
        self.weights = Array.from_numpy(np.stack((w.real, w.imag), axis=-1), np.float64,
                                        solver.device)
        self.stage = None
        self._colocated = None
        self.phasor = Array.zeros((self.nf, 6, self.n, 2), np.float64, solver.device)
        self.shape_of_grid = solver.shape
        self.data = solver.data
        widths = []
        for a, axis in enumerate("xyz"):
            lo, hi = obj["grid_slice"][a]
            v = (np.diff(solver.data[axis + "_edges_m"])[lo:hi].astype(np.float64)
                 if a != self.axis else np.ones(hi - lo))
            sh = [1, 1, 1]
            sh[a] = len(v)
            widths.append(v.reshape(sh))
        self.area_numpy = np.broadcast_to(widths[0] * widths[1] * widths[2], self.shape).copy()

    def active_steps(self, steps):
        return np.ones(steps, bool)

    def footprint(self):
        return colocation_footprint(self.obj["grid_slice"], self.shape_of_grid)

    def stage_box(self):
        if self.exact:
            return self.footprint()
        (x0, x1), (y0, y1), (z0, z1) = self.obj["grid_slice"]
        return (x0, x1, y0, y1, z0, z1)

    def replay_desc(self):
        d = ReplayDesc()
        d.stage = ptr(self.stage)
        for a, v in enumerate(self.stage_box()):
            d.box[a] = int(v)
        d.x0, d.y0, d.z0 = self.lo
        d.sx, d.sy, d.sz = self.shape
        d.kind = 2
        d.exact, d.axis = int(self.exact), int(self.axis)
        d.steps = int(self.steps)
        d.nf = int(self.nf)
        d.weights64 = ptr(self.weights)
        d.phasor64 = ptr(self.phasor)
        return d

    def reset(self):
        self.phasor.zero_()

    def launch(self, s):
        """Per-step recording (used when the solver does not stage, e.g. the
        packed-CPML kernels): colocate into a one-slot buffer, accumulate."""
        if self._colocated is None:
            self._colocated = Array.zeros((1, 6, self.n), device=s.device)
        call("wfd_dft_stage", ptr(s.E), ptr(s.H), ptr(s.hprev), ptr(s.record_step),
             ptr(self._colocated), 1, *s.shape, *self.lo, *self.shape, int(self.exact),
             stream())
        call("wfd_dft_accumulate", ptr(self._colocated), 1, ptr(self.weights), self.steps,
             self.nf, ptr(s.record_step), 1, ptr(self.phasor), self.n, stream())

    def phasors(self):
        raw = self.phasor.numpy()
        return (raw[..., 0] + 1j * raw[..., 1]).reshape(self.nf, 6, *self.shape)

    def flux(self):
        """0.5 Re sum((E x H*)[axis] * area) per wavelength (reference_report.py:89-90)."""
        ph = self.phasors()
        s = np.cross(ph[:, :3], ph[:, 3:].conj(), axis=1)[:, self.axis]
        return 0.5 * np.real(np.sum(s * self.area_numpy[None], axis=tuple(range(1, s.ndim))))

    def mode_amplitudes(self):
        """Overlap with the stored mode of every wavelength, the reference's
        alpha = sum((E_m x H* + E* x H_m)[axis])/4 (detectors.py:203-211)."""
        ph = self.phasors()
        em = self.data[self.name + "/_mode_E"]
        hm = self.data[self.name + "/_mode_H"]
        integrand = (np.cross(em, ph[:, 3:].conj(), axis=1)
                     + np.cross(ph[:, :3].conj(), hm, axis=1))[:, self.axis]
        return np.sum(integrand, axis=tuple(range(1, integrand.ndim))) / 4

    def power(self, s=None):
        return np.abs(self.mode_amplitudes()) ** 2


def record_all(solver, monitors):
    """All monitors in as few launches as possible (16 per launch)."""
    for start in range(0, len(monitors), WFD_MAX_MONITORS):
        group = monitors[start:start + WFD_MAX_MONITORS]
        descs = (MonitorDesc * len(group))(*[m.descriptor() for m in group])
        call("wfd_record_monitors", ptr(solver.E), ptr(solver.H), ptr(solver.hprev),
             ptr(solver.record_step), descs, len(group), *solver.shape, stream())


def attach(solver):
    """detectors.py:214-220"""
    for obj in solver.meta["objects"]:
        if obj["class"].endswith("Detector"):
            cls = DFTMonitor if obj["class"] == "DFTPlaneDetector" else Monitor
            m = cls(solver, obj)
            solver.monitors.append(m)
            if m.exact:
                solver.save_h_from = min(solver.save_h_from, m.first)


def measurements(solver):
    """detectors.py:223-232"""
    powers = {m.name: m.power(solver) for m in solver.monitors
              if m.obj["class"] == "ModeOverlapDetector"}
    if powers:
        powers["modal_transmission"] = powers["mode_out"] / powers["mode_in"]
        powers["modal_reflection"] = powers["mode_reflected"] / powers["mode_in"]
    return powers


def dft_transmission(solver, inlet="mode_in", outlet="mode_out"):
    """Modal transmission per wavelength of two DFT planes (no reference
    counterpart): P_out / P_in with P = |alpha|^2 of DFTMonitor.power."""
    planes = {m.name: m for m in solver.monitors if isinstance(m, DFTMonitor)}
    if inlet not in planes or outlet not in planes:
        raise ValueError(f"DFT planes {inlet!r} and {outlet!r} are required")
    a, b = planes[inlet], planes[outlet]
    if not np.allclose(a.wavelengths, b.wavelengths, rtol=1e-12, atol=0):
        raise ValueError("the two planes record different wavelengths")
    return b.power() / a.power()
