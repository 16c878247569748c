"""Fixed-input 3-D FDTD solver on the HIP kernels of libwfdtd (gfx950).

Drop-in for warp_fdtd/solver.py: same constructor, validation errors,
attributes and methods.  NumPy prepares the immutable tables on the host; the
time loop runs as the fused E+H kernel pair (cpml_storage="full") or the
per-cell packed-CPML kernels (cpml_storage="packed"), replayed through HIP
graphs captured by this library, or as the on-chip-resident cooperative kernel
(wfd_resident_run) when the whole grid fits the LDS of the co-resident
workgroups and every detector is staged.
"""

from pathlib import Path
import ctypes
import hashlib
import json
import time

import numpy as np
import torch

from . import kernels as k
from ._lib import (Handles, ReplayDesc, ResidentLayout, WFD_MAX_STAGES, YeePlan, YeeTargets, call,
                   ptr, release, stream)
from .array import Array


def pack_coefficients(coeff, shape):
    """solver.py:16-29 — rank active planes per axis; pad the six histories to
    the largest slab.  Column 3 holds the small exact integer plane rank."""
    packed = np.zeros((*coeff.shape[:2], 4), np.float32)
    packed[..., :3] = coeff
    slab_sizes = []
    for axis, length in enumerate(shape):
        active = coeff[axis, :length, 0] != 0
        packed[axis, :length, 3] = np.maximum(np.cumsum(active) - 1, 0)
        slab_sizes.append(int(active.sum()) * (int(np.prod(shape)) // length))
    return packed, max(1, *slab_sizes)


def owned_array_bytes(solver):
    """solver.py:32-50 — live device arrays owned by the solver and monitors."""
    seen = set()

    def size(value):
        if isinstance(value, Array):
            key = value.tensor.untyped_storage().data_ptr()
            if key in seen:
                return 0
            seen.add(key)
            return value.tensor.untyped_storage().nbytes()
        if isinstance(value, torch.Tensor):
            key = value.untyped_storage().data_ptr()
            if key in seen:
                return 0
            seen.add(key)
            return value.untyped_storage().nbytes()
        if isinstance(value, dict):
            return sum(size(v) for v in value.values())
        if isinstance(value, (list, tuple)):
            return sum(size(v) for v in value)
        return 0

    return size(vars(solver)) + sum(size(vars(m)) for m in solver.monitors)


def load_case(case):
    """Manifest and inputs of a bundle directory, with the sha256 check."""
    path = Path(case)
    meta = json.loads((path / "manifest.json").read_text())
    raw = path / "inputs.npz"
    if hashlib.sha256(raw.read_bytes()).hexdigest() != meta["inputs_sha256"]:
        raise ValueError("Input bundle checksum mismatch")
    with np.load(raw) as data:
        return path, meta, {n: data[n] for n in data.files}


def cpml_rows(meta, data, shape, field):
    """solver.py:105-123 — (3, max(shape), 3) rows (a, b, 1/kappa) per axis."""
    coeff = np.zeros((3, max(shape), 3), np.float32)
    coeff[:, :, 1:] = 1
    covered = np.zeros((3, max(shape)), bool)
    for obj in meta["objects"]:
        if obj["class"] != "PerfectlyMatchedLayer":
            continue
        name = obj["name"]
        dims = data[f"{name}/pml_a_{field}"].shape
        axis = int(np.argmax(dims))
        lo, hi = obj["grid_slice"][axis]
        if covered[axis, lo:hi].any():
            raise ValueError("Overlapping CPML slabs unsupported")
        covered[axis, lo:hi] = True
        for j, key in enumerate((f"pml_a_{field}", f"pml_b_{field}", f"inv_kappa_{field}")):
            coeff[axis, lo:hi, j] = data[f"{name}/{key}"].ravel()
    return coeff


PULSE_PROFILES = ("GaussianPulseProfile", "HannPulseProfile")


def pulse_waves(meta, src, steps, offsets, half):
    """Pulse wave tables (no reference counterpart; same time sampling as the
    single-frequency table): Gaussian exp(-(t-t0)^2/2s^2) cos(w (t-t0) + phi)
    or Hann sin^2(pi t/T) cos(w t + phi) on 0 <= t <= T, in float64, stored
    as float32."""
    t = (np.arange(steps, dtype=np.float64)[None, :] + half + offsets[:, :1]) * meta["dt_s"]
    w = 2 * np.pi * src["frequency_hz"]
    phi = src["phase_shift"] + src["profile_phase_shift"]
    if src["temporal_profile"] == "GaussianPulseProfile":
        u = t - src["delay_s"]
        wave = np.exp(-0.5 * (u / src["sigma_s"]) ** 2) * np.cos(w * u + phi)
    else:
        T = src["duration_s"]
        inside = (t >= 0) & (t <= T)
        wave = np.where(inside, np.sin(np.pi * t / T) ** 2, 0.0) * np.cos(w * t + phi)
    return (wave * src["static_amplitude_factor"]).astype(np.float32)


def source_waves(meta, data, src, steps, field):
    """solver.py:162-181 — float32 wave table (3, steps) of one source field,
    with the reference's separate float32 multiply/divide."""
    name = src["name"]
    offsets = data[f"{name}/_time_offset_{field}"].reshape(3, -1)
    if not np.all(offsets == offsets[:, :1]):
        raise ValueError("Spatially varying source delays unsupported")
    if src["temporal_profile"] in PULSE_PROFILES:
        return pulse_waves(meta, src, steps, offsets, 0.5 if field == "E" else 0.0)
    half_step = np.float32(0.5 if field == "E" else 0.0)
    t = (np.arange(steps, dtype=np.float32)[None, :] + half_step + offsets[:, :1]) * np.float32(
        meta["dt_s"])
    period = np.float32(1 / src["frequency_hz"])
    phase = (np.float32(2 * np.pi) * t / period + np.float32(src["phase_shift"])
             + np.float32(src["profile_phase_shift"]))
    ramp = np.clip(t / np.float32(src["num_startup_periods"] / src["frequency_hz"]), 0, 1)
    return (ramp * np.cos(phase) * np.float32(src["static_amplitude_factor"])).astype(np.float32)


class Solver:
    """solver.py:53-264 on libwfdtd.

    Extra keyword ``full_hprev``: hprev (H at the start of the step, saved once
    t >= save_h_from) is kept for every cell, as in the reference; by default
    the fused path saves it only inside the footprint the attached monitors
    read (their planes grown by the colocation stencil), which is all the
    detectors use.  VolumeMonitor or any other monitor type forces the full
    volume.

    Extra keyword ``resident``: "auto" (default) lets ``compile`` time the
    on-chip-resident stepping against the fused kernel pair and keep the
    faster where it applies (full CPML storage, every monitor staged, no
    hprev); True requires it (ValueError where it does not apply); False
    never uses it.  Both give bitwise the same fields and detectors.

    Extra keyword ``single_pass``: "auto" (default) lets ``compile`` time the
    single-pass step (wfd_yee_pass: one launch per step, E, H and CPML psi
    alternating between two buffer sets) against the fused pair where it
    applies (full CPML storage, every monitor staged) and the resident path
    is not used; True requires it; False never uses it.  Captured chunks
    pair the passes so that the state is back in ``E``, ``H``, ``psi_e``,
    ``psi_h`` after every chunk (an odd chunk ends with one fused step);
    single ``launch_step`` calls use the fused pair.  Bitwise the same
    results.
    """

    def __init__(self, case, device="cuda:0", graph_steps=64, cpml_storage="full",
                 full_hprev=False, x_chunk=0, resident="auto", single_pass="auto"):
        if cpml_storage not in ("full", "packed"):
            raise ValueError("cpml_storage must be full or packed")
        if resident not in ("auto", True, False):
            raise ValueError('resident must be "auto", True or False')
        if single_pass not in ("auto", True, False):
            raise ValueError('single_pass must be "auto", True or False')
        self.cpml_storage = cpml_storage
        self.path, self.meta, self.data = load_case(case)
        d, m = self.data, self.meta
        eps = d["inv_permittivity"]
        if (eps.ndim != 4 or eps.shape[0] != 1 or not np.all(np.isfinite(eps))
                or np.any(eps <= 0)):
            raise ValueError("Only finite positive isotropic permittivity is supported")
        if d["inv_permeability"].ndim != 0 or d["inv_permeability"] != 1:
            raise ValueError("Only relative permeability 1 is supported")
        if m["settings"].get("precision", "float32") != "float32":
            raise ValueError("Only float32 field bundles supported")
        dx = m["settings"]["spacing_nm"] * 1e-9
        for axis in "xyz":
            if not np.allclose(np.diff(d[f"{axis}_edges_m"]), dx, rtol=1e-4, atol=0):
                raise ValueError("Only uniform cubic grids supported")
        self.shape = tuple(int(v) for v in eps.shape[1:])
        self.n = int(np.prod(self.shape))
        self.steps = int(m["time_steps"])
        if self.steps < 1 or not np.isfinite(m["dt_s"]) or m["dt_s"] <= 0:
            raise ValueError("Positive time step and trajectory length required")
        factor = m["settings"].get("courant_factor", 0.99)
        if not np.isfinite(factor) or not 0 < factor <= 0.99:
            raise ValueError("Require 0 < exported CFL factor <= 0.99")
        expected_dt = factor / np.sqrt(3) * dx / 299792458.0
        if not np.isclose(m["dt_s"], expected_dt, rtol=1e-4, atol=0):
            raise ValueError("Exported time step disagrees with the declared CFL factor")
        if graph_steps < 1:
            raise ValueError("graph_steps must be positive")
        if self.n >= 2 ** 31:
            raise ValueError("Grids of 2^31 cells or more are unsupported")
        if not torch.cuda.is_available():
            raise RuntimeError("libwfdtd needs a HIP device (torch.cuda is unavailable)")
        self.device = torch.device(device)
        self.graph_steps = graph_steps
        self.courant = factor / np.sqrt(3)
        self.full_hprev = bool(full_hprev)
        self.x_chunk = int(x_chunk)
        self.E = Array.zeros((3, self.n), device=self.device)
        self.H = Array.zeros((3, self.n), device=self.device)
        self.hprev = Array.zeros((3, self.n), device=self.device)
        self.eps = Array.from_numpy(eps.reshape(-1), np.float32, self.device)
        # counters[0] = completed steps (the reference's `step`); counters[1] = the
        # step the H half-step is applying, read by monitors after the step.
        self.counters = Array.zeros(2, np.int32, self.device)
        self.step = self.counters[0:1]
        self.coeffs = {}
        self._coeff_rows = {}
        for field in "EH":
            coeff = cpml_rows(m, d, self.shape, field)
            self._coeff_rows[field] = coeff
            if cpml_storage == "packed":
                packed, psi_size = pack_coefficients(coeff, self.shape)
                self.coeffs[field] = Array.from_numpy(packed, np.float32, self.device)
            else:
                psi_size = self.n
                self.coeffs[field] = Array.from_numpy(coeff, np.float32, self.device)
            setattr(self, "psi_" + field.lower(), Array.zeros((6, psi_size), device=self.device))
        sources = [o for o in m["objects"] if o["class"] == "ModePlaneSource"]
        if len(sources) != 1:
            raise ValueError("Exactly one exported plane mode source required")
        src = sources[0]
        if (src["direction"] != "+" or src["propagation_axis"] != 0
                or src["temporal_profile"] not in ("SingleFrequencyProfile", *PULSE_PROFILES)):
            raise ValueError("Only +x single-frequency or pulsed source supported")
        self.source_x = src["grid_slice"][0][0]
        if src["grid_slice"] != [[self.source_x, self.source_x + 1], [0, self.shape[1]],
                                 [0, self.shape[2]]]:
            raise ValueError("Source must span complete yz plane")
        name = src["name"]
        if not np.array_equal(d[name + "/_time_step_to_on_idx"], np.arange(self.steps)):
            raise ValueError("Switched source schedules are not supported")
        if src["temporal_profile"] == "SingleFrequencyProfile" and src["num_startup_periods"] <= 0:
            raise ValueError("Source must have a positive startup ramp")
        self.sources, self.waves = {}, {}
        for field in "EH":
            profile = d[f"{name}/_{field}"]
            if np.iscomplexobj(profile):
                raise ValueError("Complex source profiles unsupported")
            self.sources[field] = Array.from_numpy(profile.reshape(3, -1), np.float32, self.device)
            self.waves[field] = Array.from_numpy(
                source_waves(m, d, src, self.steps, field), np.float32, self.device)
        self.monitors = []
        self.save_h_from = self.steps
        self.graphs = {}
        self._quiet_graphs = {}
        self._handles = Handles()
        self._flush_every = 1
        self._pending = 0
        self.resident_mode = resident
        self.resident = False
        self.layout = None
        self._workspace = None
        self.single_pass_mode = single_pass
        self.single_pass = False
        # (x planes per workgroup, slots per thread, steps per launch); 0 = chosen
        self.pass_shape = (0, 0, 1, 0)
        self._alt = None          # the second buffer set of the single pass
        self._ticket = None
        self._phase = 0           # 1 while the state is in the second set (inside a capture)
        self._schedule = []       # launch kind of each step left in the capture
        self._curl = None         # (history, slots, box) a two-step pass records
        self._pass_tables = {}    # device argument tables of two-step passes (batch of one)
        self.pass_table = False
        self._compiled = False

    # ------------------------------------------------------------ stepping
    @property
    def record_step(self):
        """Device counter holding the step just applied, as monitors read it."""
        return self.counters[1:2] if self.cpml_storage == "full" else self.counters[0:1]

    def staged(self):
        """Monitors recorded through footprint staging + chunk replay (fused
        path only; at most WFD_MAX_STAGES); the rest record every step."""
        if self.cpml_storage != "full":
            return []
        return [m for m in self.monitors if hasattr(m, "replay_desc")][:WFD_MAX_STAGES]

    def _stage_buffers(self):
        """One stage buffer per distinct stage box, {box: buffer}: monitors on
        the same footprint (a modal and a flux plane on one plane) share it,
        so the step stages those cells once and each replay reads them."""
        shared = {}
        for m in self.staged():
            box = tuple(int(v) for v in m.stage_box())
            if box in shared:
                m.stage = shared[box]
                continue
            count = (box[1] - box[0]) * (box[3] - box[2]) * (box[5] - box[4])
            buf = getattr(m, "stage", None)
            if buf is None or buf.shape != (self.graph_steps, 9, count) or any(
                    buf is b for b in shared.values()):
                m.stage = Array.zeros((self.graph_steps, 9, count), device=self.device)
            shared[box] = m.stage
        return shared

    def hprev_box(self):
        """Half-open (x0, x1, y0, y1, z0, z1) of cells whose hprev is read by
        monitors recorded every step: the union of their footprints (one
        without a ``footprint`` reads the whole volume).  Staged monitors take
        H_old from their stage buffers and need none."""
        nx, ny, nz = self.shape
        full = (0, nx, 0, ny, 0, nz)
        if self.full_hprev:
            return full
        staged = set(map(id, self.staged()))
        lo = [nx, ny, nz]
        hi = [0, 0, 0]
        for m in self.monitors:
            if id(m) in staged or not getattr(m, "exact", True):
                continue
            box = m.footprint() if hasattr(m, "footprint") else None
            if box is None:
                return full
            for a in range(3):
                lo[a] = min(lo[a], box[2 * a])
                hi[a] = max(hi[a], box[2 * a + 1])
        if hi[0] <= lo[0]:
            return (0, 0, 0, 0, 0, 0)
        return (lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])

    def _plan(self, stage=True):
        p = YeePlan()
        p.e, p.h = ptr(self.E), ptr(self.H)
        p.inv_eps = ptr(self.eps)
        p.psi_e, p.psi_h = ptr(self.psi_e), ptr(self.psi_h)
        p.coeff_e, p.coeff_h = ptr(self.coeffs["E"]), ptr(self.coeffs["H"])
        p.src_e, p.src_h = ptr(self.sources["E"]), ptr(self.sources["H"])
        p.wave_e, p.wave_h = ptr(self.waves["E"]), ptr(self.waves["H"])
        p.counters = ptr(self.counters)
        p.hprev = ptr(self.hprev)
        for a, v in enumerate(self.hprev_box()):
            p.hbox[a] = int(v)
        p.save_h_from = int(self.save_h_from)
        p.coeff_rows = int(self.coeffs["E"].shape[1])
        p.steps = int(self.steps)
        p.nx, p.ny, p.nz = self.shape
        p.source_x = int(self.source_x)
        p.courant = float(self.courant)
        p.x_chunk = self.x_chunk
        p.slots = int(self.graph_steps)
        if stage:
            shared = self._stage_buffers()
            p.nstage = len(shared)
            for b, (box, buf) in enumerate(shared.items()):
                p.stages[b].buf = ptr(buf)
                for a, v in enumerate(box):
                    p.stages[b].box[a] = v
        return p

    def reset(self):
        for a in (self.E, self.H, self.hprev, self.psi_e, self.psi_h, self.counters):
            a.zero_()
        for monitor in self.monitors:
            monitor.reset()
        self._pending = 0
        torch.cuda.synchronize(self.device)

    def _step_fields(self, stage=True):
        if self.cpml_storage == "full":
            if self._schedule:
                self._pass_step(stage)
                return
            k.yee_step(self._plan(stage))
            return
        common = [self.hprev, self.save_h_from, *self.shape, self.source_x, self.courant, 0]
        k.update_e_packed(self.E, self.H, self.eps, self.psi_e, self.coeffs["E"],
                          self.sources["H"], self.waves["H"], self.step, *common)
        k.update_h_packed(self.H, self.E, self.eps, self.psi_h, self.coeffs["H"],
                          self.sources["E"], self.waves["E"], self.step, *common)

    def _replay_staged(self, m_steps):
        staged = self.staged()
        if not staged or m_steps == 0:
            return
        descs = (ReplayDesc * len(staged))(*[m.replay_desc() for m in staged])
        call("wfd_replay_monitors", descs, len(staged), int(self.graph_steps),
             ptr(self.record_step), int(m_steps), *self.shape, stream())

    def _record(self):
        """Monitors recorded every step; staged ones replay once per chunk."""
        from .detectors import Monitor, record_all

        staged = set(map(id, self.staged()))
        plain = [m for m in self.monitors if isinstance(m, Monitor) and id(m) not in staged]
        if plain:
            record_all(self, plain)
        for m in self.monitors:
            if not isinstance(m, Monitor) and id(m) not in staged:
                m.launch(self)
        if staged:
            self._pending += 1
            if self._pending == self._flush_every:
                self._replay_staged(self._pending)
                self._pending = 0

    def _finish_step(self):
        if self.cpml_storage != "full":
            k.advance(self.step)

    # ------------------------------------------------------ single-pass path
    def single_pass_applies(self):
        """Whether wfd_yee_pass can step this solver: full CPML storage and
        every monitor staged (a monitor recorded every step would read the
        buffer set the step just left), nz below 1024."""
        return (self.cpml_storage == "full" and len(self.staged()) == len(self.monitors)
                and self.shape[2] < 1024)

    def use_single_pass(self, planes=0, slots_per_thread=0, steps=1, threads=0, table=True):
        """Step captured chunks with wfd_yee_pass, ``steps`` (1 or 2) steps per
        launch, ``threads`` per workgroup (0: the library's choice for the
        shape; allocates the second buffer set).  ``table``: two-step passes
        of one slot per 1024-thread workgroup launch as a batch of one
        (wfd_batch_pass, arguments in a device table)."""
        if not self.single_pass_applies():
            raise ValueError("the single-pass step needs full CPML storage, staged monitors "
                             "only and nz < 1024")
        if steps not in (1, 2):
            raise ValueError("the single pass applies 1 or 2 steps per launch")
        if self._alt is None:
            self._alt = {"E": Array.zeros((3, self.n), device=self.device),
                         "H": Array.zeros((3, self.n), device=self.device),
                         "psi_e": Array.zeros(self.psi_e.shape, device=self.device),
                         "psi_h": Array.zeros(self.psi_h.shape, device=self.device)}
            self._ticket = Array.zeros(1, np.int32, self.device)
        self.pass_shape = (int(planes), int(slots_per_thread), int(steps), int(threads))
        self.pass_table = bool(table)
        self.single_pass = True
        self._prepare_pass_tables()

    def _pass_schedule(self, length, two_step=True):
        """Launch kind of each step of a captured chunk of ``length`` steps:
        "p2" (a two-step launch) then "skip" (its second step), "p1" (one
        step) or "fused" (wfd_yee_step in place).  Passes come in pairs, so
        the state is back in the first set after every pair; two-step pairs
        cover 4 steps, and chunk flushes of the detector stages (every
        graph_steps steps) never fall inside a two-step launch."""
        two = two_step and self.pass_shape[2] == 2 and self.graph_steps % 2 == 0
        kinds, left = [], int(length)
        if two:
            while left >= 4:
                kinds += ["p2", "skip", "p2", "skip"]
                left -= 4
        while left >= 2:
            kinds += ["p1", "p1"]
            left -= 2
        return kinds + ["fused"] * left

    def _pass_step(self, stage):
        """The next step of the capture's schedule: a pass from the current
        set to the other (one or two steps; wfd_yee_pass, or wfd_batch_pass
        with the prepared table of these arguments), nothing (the second step
        of a two-step launch) or the fused pair in place."""
        kind = self._schedule.pop(0)
        if kind == "skip":
            return
        if kind == "fused":
            if self._phase:
                raise RuntimeError("a fused step with the state in the second set")
            k.yee_step(self._plan(stage))
            return
        plan, out = self._pass_args(stage, self._phase, 2 if kind == "p2" else 1, self._curl)
        table = self._pass_tables.get(bytes(plan) + bytes(out))
        if table is not None:
            call("wfd_batch_pass", ctypes.byref(table[0]), ctypes.byref(table[1]), 1,
                 ptr(table[2]), stream())
        else:
            call("wfd_yee_pass", ctypes.byref(plan), ctypes.byref(out), stream())
        self._phase ^= 1

    def _pass_args(self, stage, phase, steps, curl):
        """(plan, targets) of a pass from buffer set ``phase`` to the other."""
        main = {"E": self.E, "H": self.H, "psi_e": self.psi_e, "psi_h": self.psi_h}
        src, dst = (main, self._alt) if phase == 0 else (self._alt, main)
        plan = self._plan(stage)
        plan.e, plan.h = ptr(src["E"]), ptr(src["H"])
        plan.psi_e, plan.psi_h = ptr(src["psi_e"]), ptr(src["psi_h"])
        out = YeeTargets()
        out.e, out.h = ptr(dst["E"]), ptr(dst["H"])
        out.psi_e, out.psi_h = ptr(dst["psi_e"]), ptr(dst["psi_h"])
        out.ticket = ptr(self._ticket)
        out.steps = steps
        if out.steps == self.pass_shape[2]:
            out.planes, out.slots_per_thread = self.pass_shape[:2]
            out.threads = self.pass_shape[3]
        if steps == 2 and curl is not None:
            hist, slots, box = curl
            out.hist, out.hist_steps = ptr(hist), int(slots)
            for q, v in enumerate(box):
                out.hist_box[q] = int(v)
        return plan, out

    def _prepare_pass_tables(self):
        """With ``pass_table``, two-step passes of one slot per thread in
        1024-thread workgroups launch as a batch of one (wfd_batch_pass): the
        kernel reads its arguments from a device table instead of the kernarg
        segment and keeps them out of the scalar registers it would otherwise
        spill.  The tables (per stage flag, buffer set and curl target) are
        written here, outside any capture; a pass whose arguments match none
        takes wfd_yee_pass."""
        self._pass_tables = {}
        planes, slots, steps, threads = self.pass_shape
        if not self.pass_table or steps != 2 or slots not in (0, 1) or threads not in (0, 1024):
            return
        from ._lib import lib

        nbytes = int(lib().wfd_batch_pass_workspace_bytes(1))
        curls = [None] + ([self.history_target()] if hasattr(self, "history_target") else [])
        for stage in (True, False):
            for phase in (0, 1):
                for curl in curls:
                    plan, out = self._pass_args(stage, phase, 2, curl)
                    ws = Array.zeros((nbytes + 3) // 4, np.int32, self.device)
                    call("wfd_batch_pass_prepare", ctypes.byref(plan), ctypes.byref(out), 1,
                         ptr(ws), nbytes, stream())
                    self._pass_tables[bytes(plan) + bytes(out)] = (plan, out, ws)

    def current_h(self):
        """H of the state the next step starts from (the second set while a
        captured chunk is between two passes)."""
        return self.H if self._phase == 0 else self._alt["H"]

    # ------------------------------------------------- on-chip-resident path
    def resident_applies(self):
        """Whether the resident kernel can run this solver's step: full CPML
        storage, every monitor recorded through footprint staging and no hprev
        to save (it keeps no H_old outside the stage buffers)."""
        if self.cpml_storage != "full" or self.full_hprev:
            return False
        if len(self.staged()) != len(self.monitors):
            return False
        box = self.hprev_box()
        return box[1] <= box[0]

    def resident_layout(self, sx=0, sy=0, sz=0):
        """wfd_resident_layout_for: the sub-box decomposition (ValueError when
        the grid does not fit the co-resident workgroups)."""
        layout = ResidentLayout()
        call("wfd_resident_layout_for", ctypes.byref(self._plan(stage=False)), int(sx), int(sy),
             int(sz), ctypes.byref(layout))
        return layout

    def use_resident(self, layout=None):
        """Switch stepping to the resident kernel with the given (default:
        chosen) layout; its workspace starts zeroed (the face words carry
        launch-epoch tags and need no clearing between launches of one
        layout, only when the layout using them changes)."""
        if not self.resident_applies():
            raise ValueError("resident stepping needs full CPML storage, staged monitors "
                             "only and no hprev")
        self.layout = layout if layout is not None else self.resident_layout()
        words = int(self.layout.workspace_bytes) // 4
        if self._workspace is None or self._workspace.shape[0] != words:
            self._workspace = Array.zeros(words, np.int32, self.device)
        else:
            self._workspace.zero_()
        self.resident = True

    def _error_word(self):
        """Index (int32) of the error word: after the 64-bit face words."""
        L = self.layout
        return 2 * 4 * (L.sy * L.sz + L.sx * L.sz + L.sx * L.sy) * L.blocks

    def check(self):
        """Raise if a resident run timed out waiting for a neighbouring
        workgroup (its fields are then invalid); synchronizes."""
        torch.cuda.synchronize(self.device)
        if not self.resident:
            return
        i = self._error_word()
        if int(self._workspace[i:i + 1].numpy()[0]) != 0:
            self._workspace.zero_()
            raise RuntimeError("resident stepping timed out waiting for a neighbouring "
                               "workgroup; the fields of this run are invalid")

    def resident_steps(self, count, record=True, spin_limit=0, events=None):
        """`count` steps as one cooperative launch; staged monitors (record)
        replay afterwards, so count must not exceed graph_steps then.
        `events` (a list): a (start, end, count) tuple of timing events
        recorded on the launch stream around the launch is appended."""
        if record and self.monitors and count > self.graph_steps:
            raise ValueError("recorded resident chunks are at most graph_steps long")
        plan = self._plan(stage=record)
        if events is not None:
            pair = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            pair[0].record()
        call("wfd_resident_run", ctypes.byref(plan), ctypes.byref(self.layout),
             ptr(self._workspace), int(count), int(spin_limit), stream())
        if events is not None:
            pair[1].record()
            events.append((pair[0], pair[1], int(count)))
        if record:
            self._replay_staged(count)

    def launch_step(self, record=True):
        """solver.py:194-237 — E, H, monitors, time advance."""
        self._step_fields(stage=record)
        if record:
            self._record()
        self._finish_step()

    def _capture_fn(self, length, fn, two_step=True, stepping=True):
        """HIP graph of `length` calls of fn, captured on a library stream;
        staged monitors replay every graph_steps steps and at the end.
        ``stepping``: each call of fn takes one step (through _step_fields);
        ``two_step``: the chunk may use two-step passes (not when fn needs
        the fields between the two steps)."""
        torch.cuda.synchronize(self.device)
        # the stage rings hold graph_steps steps: longer graphs flush inside
        self._flush_every, self._pending = min(length, self.graph_steps), 0
        self._phase = 0
        self._schedule = (self._pass_schedule(length, two_step)
                          if self.single_pass and stepping else [])
        s = self._handles.begin()
        ext = torch.cuda.ExternalStream(s.value, device=self.device)
        try:
            with torch.cuda.stream(ext):
                for _ in range(length):
                    fn()
                if self._pending:
                    self._replay_staged(self._pending)
        finally:
            graph = self._handles.end()
            self._flush_every, self._pending = 1, 0
            ended_in, left = self._phase, len(self._schedule)
            self._phase, self._schedule = 0, []
        if ended_in or left:
            raise RuntimeError("a captured chunk ended with the state in the second set")
        return graph

    def _capture(self, length, record):
        return self._capture_fn(length, lambda: self.launch_step(record=record))

    def tune(self):
        """Pick the fused step's launch shape for this grid and, with
        resident="auto", whether the resident kernel is faster (autotune.py)."""
        from .autotune import tune, tune_resident

        if self.cpml_storage == "full" and self.x_chunk == 0:
            self.x_chunk = tune(self)
        self.resident = False
        self.single_pass = False
        if self.resident_mode is True:
            if not self.resident_applies():
                raise ValueError("resident stepping needs full CPML storage, staged monitors "
                                 "only and no hprev")
            self.use_resident(tune_resident(self, require=True))
        elif self.resident_mode == "auto" and self.resident_applies():
            layout = tune_resident(self)
            if layout is not None:
                self.use_resident(layout)
        if self.resident:
            return self.x_chunk
        from .autotune import tune_pass

        if self.single_pass_mode is True:
            self.use_single_pass(*tune_pass(self, require=True))
        elif self.single_pass_mode == "auto" and self.single_pass_applies():
            shape = tune_pass(self)
            if shape is not None:
                self.use_single_pass(*shape)
        return self.x_chunk

    def compile(self):
        """solver.py:239-250 — warm step, reset, capture graph_steps and tail."""
        start = time.perf_counter()
        self.launch_step()
        torch.cuda.synchronize(self.device)
        self.tune()
        self.capture_all()
        return time.perf_counter() - start

    def capture_all(self):
        """(Re)capture the step graphs with the current stepping choices
        (compile's tail: after tune, or after use_single_pass/use_resident)."""
        self.reset()
        self._destroy_graphs()
        if not self.resident:
            for length in {self.graph_steps, self.steps % self.graph_steps} - {0}:
                self.graphs[length] = self._capture(length, True)
                if self.monitors:
                    self._quiet_graphs[length] = self._capture(length, False)
        self.reset()
        self._compiled = True

    def _windows(self):
        """Boolean per step: does any monitor record at that step?"""
        active = np.zeros(self.steps, bool)
        for m in self.monitors:
            active |= m.active_steps(self.steps)
        return active

    def _chunks(self):
        return [self.graph_steps] * (self.steps // self.graph_steps) + (
            [self.steps % self.graph_steps] if self.steps % self.graph_steps else [])

    def _run_resident(self, events=None):
        """Chunks with an active detector window run as recorded launches of
        graph_steps; runs of quiet chunks merge into one launch."""
        active = self._windows()
        t, quiet = 0, 0
        for length in self._chunks():
            if self.monitors and active[t:t + length].any():
                if quiet:
                    self.resident_steps(quiet, record=False, events=events)
                    quiet = 0
                self.resident_steps(length, record=True, events=events)
            else:
                quiet += length
            t += length
        if quiet:
            self.resident_steps(quiet, record=False, events=events)
        self.check()

    def run(self, events=None):
        """solver.py:252-258 — replay the captured graphs over all steps.
        `events` (a list, resident stepping only): see resident_steps."""
        if not self._compiled:
            raise RuntimeError("compile() must run before run()")
        if self.resident:
            self._run_resident(events)
            return
        active = self._windows()
        s = stream()
        t = 0
        for length in self._chunks():
            quiet = self._quiet_graphs.get(length)
            graph = quiet if (quiet is not None and not active[t:t + length].any()) \
                else self.graphs[length]
            call("wfd_graph_launch", graph, s)
            t += length
        torch.cuda.synchronize(self.device)

    def run_steps(self, count, synchronize=True, events=None):
        """Advance exactly `count` steps from the current step with graph
        replays (graphs of missing lengths are captured on first use), or
        resident launches of at most graph_steps steps (`events`: see
        resident_steps)."""
        if not self._compiled:
            raise RuntimeError("compile() must run before run_steps()")
        if self.resident:
            left = int(count)
            while left > 0:
                length = min(self.graph_steps, left) if self.monitors else left
                self.resident_steps(length, record=bool(self.monitors), events=events)
                left -= length
            if synchronize:
                self.check()
            return
        s = stream()
        left = int(count)
        while left > 0:
            length = min(self.graph_steps, left)
            if length not in self.graphs:
                # capture records the launches without running them
                self.graphs[length] = self._capture(length, True)
            call("wfd_graph_launch", self.graphs[length], s)
            left -= length
        if synchronize:
            torch.cuda.synchronize(self.device)

    def fields(self):
        return {name: arr.numpy().reshape(3, *self.shape)
                for name, arr in (("E", self.E), ("H", self.H))}

    def _destroy_graphs(self):
        for table in (self.graphs, self._quiet_graphs):
            for g in table.values():
                self._handles.destroy(g)
            table.clear()

    def close(self):
        """Destroy this solver's HIP graphs and capture stream (idempotent;
        also done at interpreter exit).  The solver cannot replay graphs
        afterwards."""
        self.graphs.clear()
        self._quiet_graphs.clear()
        self._compiled = False
        self._handles.close()

    def __del__(self, _release=release):
        # (bound at definition: module globals may already be None when an
        # object is collected during interpreter shutdown)
        _release(getattr(self, "_handles", None))
