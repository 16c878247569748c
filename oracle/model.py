"""Host-side restatement of the reference solver, detectors and adjoint drivers
on top of the C++ oracle kernels (TEST INFRASTRUCTURE ONLY).

Independent of the product package: every table (CPML rows, source waveforms,
detector weights and areas, adjoint cotangents) is recomputed here from the
bundle, following the reference line by line, so that tests can compare the
product's tables and trajectories against it.
"""

import hashlib
import json
from pathlib import Path

import numpy as np

from . import kernels as K


# ---------------------------------------------------------------- bundle I/O
def load_bundle(case):
    """solver.py:60-66 — manifest + inputs.npz with the sha256 check."""
    path = Path(case)
    meta = json.loads((path / "manifest.json").read_text())
    raw = path / "inputs.npz"
    if hashlib.sha256(raw.read_bytes()).hexdigest() != meta["inputs_sha256"]:
        raise ValueError("Input bundle checksum mismatch")
    with np.load(raw) as data:
        return meta, {n: data[n] for n in data.files}


def cpml_table(meta, data, shape, field):
    """solver.py:105-123 — (3, max(shape), 3) rows (a, b, 1/kappa)."""
    coeff = np.zeros((3, max(shape), 3), np.float32)
    coeff[:, :, 1:] = 1
    for obj in meta["objects"]:
        if obj["class"] != "PerfectlyMatchedLayer":
            continue
        name = obj["name"]
        dims = data[f"{name}/pml_a_{field}"].shape
        axis = int(np.argmax(dims))
        lo, hi = obj["grid_slice"][axis]
        for j, key in enumerate((f"pml_a_{field}", f"pml_b_{field}", f"inv_kappa_{field}")):
            coeff[axis, lo:hi, j] = data[f"{name}/{key}"].ravel()
    return coeff


def pack_coefficients(coeff, shape):
    """solver.py:16-29"""
    packed = np.zeros((*coeff.shape[:2], 4), np.float32)
    packed[..., :3] = coeff
    sizes = []
    for axis, length in enumerate(shape):
        active = coeff[axis, :length, 0] != 0
        packed[axis, :length, 3] = np.maximum(np.cumsum(active) - 1, 0)
        sizes.append(int(active.sum()) * (int(np.prod(shape)) // length))
    return packed, max(1, *sizes)


def pulse_table(meta, src, steps, offsets, half):
    """Pulse sources of the synthetic workloads (no reference counterpart):
    Gaussian exp(-u^2/2s^2) cos(w u + phi), u = t - delay, or Hann
    sin^2(pi t/T) cos(w t + phi) on [0, T], evaluated in float64 on the
    reference's time sampling t = (step + half + offset) dt and stored float32."""
    t = (np.arange(steps, dtype=np.float64)[None, :] + half + offsets[:, :1]) * meta["dt_s"]
    omega = 2 * np.pi * src["frequency_hz"]
    phi = src["phase_shift"] + src["profile_phase_shift"]
    if src["temporal_profile"] == "GaussianPulseProfile":
        u = t - src["delay_s"]
        wave = np.exp(-0.5 * (u / src["sigma_s"]) ** 2) * np.cos(omega * u + phi)
    else:
        inside = (t >= 0) & (t <= src["duration_s"])
        wave = np.where(inside, np.sin(np.pi * t / src["duration_s"]) ** 2, 0.0) * np.cos(
            omega * t + phi)
    return (wave * src["static_amplitude_factor"]).astype(np.float32)


def source_tables(meta, data, steps):
    """solver.py:132-182 — profiles (3, ny*nz) and waves (3, steps), float32."""
    src = next(o for o in meta["objects"] if o["class"] == "ModePlaneSource")
    name = src["name"]
    profiles, waves = {}, {}
    for field in "EH":
        profiles[field] = data[f"{name}/_{field}"].reshape(3, -1).astype(np.float32)
        offsets = data[f"{name}/_time_offset_{field}"].reshape(3, -1)
        if src["temporal_profile"] in ("GaussianPulseProfile", "HannPulseProfile"):
            waves[field] = pulse_table(meta, src, steps, offsets, 0.5 if field == "E" else 0.0)
            continue
        half = np.float32(0.5 if field == "E" else 0.0)
        t = (np.arange(steps, dtype=np.float32)[None, :] + half + offsets[:, :1]) * np.float32(
            meta["dt_s"])
        period = np.float32(1 / src["frequency_hz"])
        phase = (np.float32(2 * np.pi) * t / period + np.float32(src["phase_shift"])
                 + np.float32(src["profile_phase_shift"]))
        ramp = np.clip(t / np.float32(src["num_startup_periods"] / src["frequency_hz"]), 0, 1)
        waves[field] = (ramp * np.cos(phase) * np.float32(src["static_amplitude_factor"])).astype(
            np.float32)
    return src["grid_slice"][0][0], profiles, waves


class Monitor:
    """detectors.py:109-211 (host part)."""

    def __init__(self, meta, data, steps, obj, dtype):
        self.obj, self.name = obj, obj["name"]
        self.is_phasor = obj["class"] in ("ModeOverlapDetector", "PhasorDetector")
        self.shape = tuple(hi - lo for lo, hi in obj["grid_slice"])
        self.lo = tuple(lo for lo, hi in obj["grid_slice"])
        self.n = int(np.prod(self.shape))
        self.exact = self.is_phasor or meta["settings"].get("exact_interpolation", False)
        self.indices = np.ascontiguousarray(data[self.name + "/_time_step_to_arr_idx"], np.int32)
        self.first = int(np.flatnonzero(self.indices >= 0)[0])
        w = np.zeros(steps, np.complex64)
        if self.is_phasor:
            src = next(o for o in meta["objects"] if o["class"] == "ModePlaneSource")
            t = np.arange(steps, dtype=np.float32) * np.float32(meta["dt_s"])
            phase = np.float32(2 * np.pi * src["frequency_hz"]) * t
            w = (np.exp(1j * phase).astype(np.complex64) * np.float32(2 / obj["window_sum"])
                 * data[self.name + "/_window_at_time_step_arr"])
        self.weights = np.ascontiguousarray(np.column_stack((w.real, w.imag)), dtype)
        self.phasor = np.zeros((6, self.n if self.is_phasor else 1, 2), dtype)
        self.flux = np.zeros(int(self.indices.max()) + 1, dtype)
        self.axis = obj.get("propagation_axis", 2)
        widths = []
        for a, axis in enumerate("xyz"):
            lo, hi = obj["grid_slice"][a]
            v = (np.diff(data[axis + "_edges_m"])[lo:hi] if a != self.axis
                 else np.ones(hi - lo, np.float32))
            sh = [1, 1, 1]
            sh[a] = len(v)
            widths.append(v.reshape(sh))
        self.area_numpy = np.broadcast_to(widths[0] * widths[1] * widths[2], self.shape).copy()
        if obj["class"] == "ModeOverlapDetector":
            self.area_numpy /= self.area_numpy.mean()
        self.area = np.ascontiguousarray(self.area_numpy.ravel(), dtype)

    def reset(self):
        self.phasor[:] = 0
        self.flux[:] = 0

    def phasors(self):
        return (self.phasor[..., 0] + 1j * self.phasor[..., 1]).reshape(6, *self.shape)

    def power(self, data):
        ph = self.phasors()
        em = data[self.name + "/_mode_E"][0]
        hm = data[self.name + "/_mode_H"][0]
        integrand = (np.cross(em, ph[3:].conj(), axis=0)
                     + np.cross(ph[:3].conj(), hm, axis=0))[self.axis]
        alpha = np.sum(integrand * self.area_numpy) / 4
        return float(abs(alpha) ** 2)


class DFTPlane:
    """Multi-wavelength running DFT plane (no reference counterpart; the
    per-wavelength restatement of detectors.py:65-106's phasor branch):
    phasor_f += colocated(E, H)(t) * exp(i 2 pi f t dt) dt, float64."""

    def __init__(self, meta, data, steps, obj):
        self.obj, self.name = obj, obj["name"]
        self.shape = tuple(hi - lo for lo, hi in obj["grid_slice"])
        self.lo = tuple(lo for lo, hi in obj["grid_slice"])
        self.n = int(np.prod(self.shape))
        self.axis = int(obj.get("propagation_axis", 0))
        self.exact = bool(obj.get("exact_interpolation", True))
        self.first = 0
        self.wavelengths = np.asarray(obj["wavelengths_m"], np.float64)
        f = 299792458.0 / self.wavelengths
        t = np.arange(steps, dtype=np.float64) * float(meta["dt_s"])
        w = np.exp(2j * np.pi * f[None, :] * t[:, None]) * float(meta["dt_s"])
        self.weights = np.ascontiguousarray(np.stack((w.real, w.imag), axis=-1))
        self.phasor = np.zeros((len(f), 6, self.n, 2))

    def reset(self):
        self.phasor[:] = 0

    def phasors(self):
        return (self.phasor[..., 0] + 1j * self.phasor[..., 1]).reshape(
            len(self.wavelengths), 6, *self.shape)

    def power(self, data, f):
        """|alpha_f|^2, alpha = sum((E_m x H* + E* x H_m)[axis]) / 4."""
        ph = self.phasors()[f]
        em = data[self.name + "/_mode_E"][f]
        hm = data[self.name + "/_mode_H"][f]
        alpha = np.sum((np.cross(em, ph[3:].conj(), axis=0)
                        + np.cross(ph[:3].conj(), hm, axis=0))[self.axis]) / 4
        return float(abs(alpha) ** 2)


class OracleSolver:
    """solver.py:53-264 restated on the CPU oracle kernels; ``dtype`` float32
    restates the reference, float64 gives the double-precision trajectory."""

    def __init__(self, case, dtype=np.float32, cpml_storage="full", history_box=None):
        self.meta, self.data = load_bundle(case)
        m, d = self.meta, self.data
        self.dtype = np.dtype(dtype)
        eps = d["inv_permittivity"]
        self.shape = tuple(int(v) for v in eps.shape[1:])
        self.n = int(np.prod(self.shape))
        self.steps = int(m["time_steps"])
        self.courant = m["settings"].get("courant_factor", 0.99) / np.sqrt(3)
        self.E = np.zeros((3, self.n), dtype)
        self.H = np.zeros((3, self.n), dtype)
        self.hprev = np.zeros((3, self.n), dtype)
        self.eps = np.ascontiguousarray(eps.reshape(-1), dtype)
        self.step = np.zeros(1, np.int32)
        self.packed = cpml_storage == "packed"
        self.coeffs, self.psi = {}, {}
        for field in "EH":
            c = cpml_table(m, d, self.shape, field)
            size = self.n
            if self.packed:
                c, size = pack_coefficients(c, self.shape)
            self.coeffs[field] = np.ascontiguousarray(c, dtype)
            self.psi[field] = np.zeros((6, size), dtype)
        self.source_x, prof, waves = source_tables(m, d, self.steps)
        self.sources = {k: np.ascontiguousarray(v, dtype) for k, v in prof.items()}
        self.waves = {k: np.ascontiguousarray(v, dtype) for k, v in waves.items()}
        self.monitors = []
        self.save_h_from = self.steps
        self.box = history_box
        self.history = None
        if history_box is not None:
            self.design_n = int(np.prod(history_box[3:]))
            self.history = np.zeros((self.steps, 3, self.design_n), dtype)

    def attach(self):
        """detectors.py:214-220"""
        for obj in self.meta["objects"]:
            if obj["class"] == "DFTPlaneDetector":
                mon = DFTPlane(self.meta, self.data, self.steps, obj)
                self.monitors.append(mon)
                if mon.exact:
                    self.save_h_from = 0
            elif obj["class"].endswith("Detector"):
                mon = Monitor(self.meta, self.data, self.steps, obj, self.dtype)
                self.monitors.append(mon)
                if mon.exact:
                    self.save_h_from = min(self.save_h_from, mon.first)

    def reset(self):
        for a in (self.E, self.H, self.hprev, self.psi["E"], self.psi["H"], self.step):
            a[...] = 0
        for mon in self.monitors:
            mon.reset()

    def launch_step(self):
        """solver.py:194-237 (and adjoint.py:271-327 with a history box)."""
        K.update(True, self.E, self.H, self.eps, self.psi["E"], self.coeffs["E"],
                 self.sources["H"], self.waves["H"], self.step, self.hprev,
                 self.save_h_from, self.shape, self.source_x, self.courant, 0, self.packed)
        if self.history is not None:
            K.record_curl(self.H, self.history, self.step, self.shape, self.box, self.courant)
        K.update(False, self.H, self.E, self.eps, self.psi["H"], self.coeffs["H"],
                 self.sources["E"], self.waves["E"], self.step, self.hprev,
                 self.save_h_from, self.shape, self.source_x, self.courant, 0, self.packed)
        for mon in self.monitors:
            if isinstance(mon, DFTPlane):
                K.record_dft(self.E, self.H, self.hprev, int(self.step[0]), mon.weights,
                             mon.phasor, self.shape, mon.lo, mon.shape, mon.exact)
                continue
            K.record(self.E, self.H, self.hprev, self.step, mon.weights, mon.indices,
                     mon.phasor, mon.flux, mon.area, self.shape, mon.lo, mon.shape,
                     mon.axis, mon.exact, mon.is_phasor)
        self.step[0] += 1

    def set_density(self, box, rho, eps_lo, eps_hi):
        """Linear density map (no reference counterpart): inv_eps = 1/(eps_lo +
        rho (eps_hi - eps_lo)) through the box's z-layers, float64 arithmetic."""
        x0, y0, z0, sx, sy, sz = box
        inv = 1 / (eps_lo + np.asarray(rho, np.float64) * (eps_hi - eps_lo))
        eps = self.eps.reshape(self.shape)
        eps[x0:x0 + sx, y0:y0 + sy, z0:z0 + sz] = inv.astype(self.dtype)[:, :, None]

    def dft_transmission(self, f=0):
        """P(mode_out) / P(mode_in) of the DFT planes at wavelength index f."""
        planes = {m.name: m for m in self.monitors if isinstance(m, DFTPlane)}
        return planes["mode_out"].power(self.data, f) / planes["mode_in"].power(self.data, f)

    def run(self, steps=None):
        for _ in range(self.steps if steps is None else steps):
            self.launch_step()

    def fields(self):
        return {"E": self.E.reshape(3, *self.shape), "H": self.H.reshape(3, *self.shape)}

    def measurements(self):
        """detectors.py:223-232"""
        powers = {m.name: m.power(self.data) for m in self.monitors
                  if m.obj["class"] == "ModeOverlapDetector"}
        if powers:
            powers["modal_transmission"] = powers["mode_out"] / powers["mode_in"]
            powers["modal_reflection"] = powers["mode_reflected"] / powers["mode_in"]
        return powers


def adjoint_sources(solver):
    """adjoint.py:329-366 — cotangent plane sources, (6, count, 2) each."""
    monitors = {m.name: m for m in solver.monitors}
    pin = monitors["mode_in"].power(solver.data)
    pout = monitors["mode_out"].power(solver.data)
    if not np.isfinite(pin) or pin <= 0 or not np.isfinite(pout):
        raise ValueError("Finite modal powers and positive incident power required")
    out = []
    for name, factor in (("mode_in", pout / pin ** 2), ("mode_out", -1 / pin)):
        m = monitors[name]
        ph = m.phasors()
        em = solver.data[name + "/_mode_E"][0]
        hm = solver.data[name + "/_mode_H"][0]
        a, b = (m.axis + 1) % 3, (m.axis + 2) % 3
        qe = np.zeros_like(em)
        qh = np.zeros_like(hm)
        qe[a] = hm[b].conj() / 4
        qe[b] = -hm[a].conj() / 4
        qh[b] = em[a].conj() / 4
        qh[a] = -em[b].conj() / 4
        coeff = np.concatenate((qe, qh)) * m.area_numpy[None]
        beta = np.sum(coeff * ph)
        source = (2 * factor * beta.conjugate() * coeff).reshape(6, -1).astype(np.complex64)
        host = np.ascontiguousarray(np.stack((source.real, source.imag), axis=-1), solver.dtype)
        out.append((m, host))
    return out


def reverse(solver, sources, grad_dtype=None):
    """adjoint.py:388-439 — full reverse sweep with the stored curl history.
    Returns the voxel gradient (design_n,) in ``grad_dtype``."""
    dt = solver.dtype
    n, shape = solver.n, solver.shape
    qe = np.zeros((3, n), dt)
    qh = np.zeros((3, n), dt)
    psie = np.zeros_like(solver.psi["E"])
    psih = np.zeros_like(solver.psi["H"])
    d = np.zeros((6, n), dt)
    grad = np.zeros(solver.design_n, grad_dtype or dt)
    step = np.array([solver.steps - 1], np.int32)
    for _ in range(solver.steps):
        for m, v in sources:
            K.inject(qe, qh, v, m.weights, step, shape, m.lo, m.shape, 0)
        K.prepare(qh, solver.eps, psih, solver.coeffs["H"], d, shape, solver.courant, 0)
        K.gather(qe, d, solver.history, grad, step, shape, solver.box, 1)
        K.prepare(qe, solver.eps, psie, solver.coeffs["E"], d, shape, solver.courant, 1)
        K.gather(qh, d, solver.history, grad, step, shape, solver.box, 0)
        for m, v in sources:
            K.inject(qe, qh, v, m.weights, step, shape, m.lo, m.shape, 1)
        step[0] -= 1
    return grad
