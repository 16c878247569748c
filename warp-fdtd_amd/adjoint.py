"""Exact discrete material adjoint on libwfdtd (drop-in for warp_fdtd/adjoint.py).

The reverse sweep is the exact transpose of the Yee + CPML step and of the
detector colocation.  Two storage schemes for the forward curl(H) the
gradient needs inside the design box:

* ``checkpoint_every=None`` — the reference's full history, (steps, 3, box)
  float32 (adjoint.py:256-260);
* ``checkpoint_every=K`` — E, H and CPML state saved every K steps during the
  forward; the reverse restores a checkpoint, recomputes its K steps (bitwise,
  the forward has no atomics) into a K-step curl buffer and walks them back.

The voxel gradient accumulates in float64 (``grad_dtype``).
"""

import ctypes
import time

import numpy as np
import torch

from . import kernels as k
from ._lib import WFD_MAX_INJECT, InjectSource, call, ptr, stream
from .array import Array
from .kernels import Kernel
from .solver import Solver


# ------------------------------------------------------------- kernel ABI
def _record_curl(h, hist, step, nx, ny, nz, x0, y0, z0, sx, sy, sz, courant, dim=None):
    call("wfd_record_curl", ptr(h), ptr(hist), int(hist.shape[0]), ptr(step), int(nx),
         int(ny), int(nz), int(x0), int(y0), int(z0), int(sx), int(sy), int(sz),
         float(courant), stream())


def _prepare(q, eps, psi, coeff, d, nx, ny, nz, courant, electric):
    call("wfd_adjoint_prepare", ptr(q), ptr(eps), ptr(psi), ptr(coeff), int(coeff.shape[1]),
         ptr(d), int(nx), int(ny), int(nz), float(courant), int(electric), stream())


def _gather(q, d, hist, grad, step, nx, ny, nz, x0, y0, z0, sx, sy, sz, from_h):
    t = getattr(grad, "tensor", grad)
    name = "wfd_adjoint_gather_f64" if t.dtype == torch.float64 else "wfd_adjoint_gather"
    call(name, ptr(q), ptr(d), ptr(hist), int(hist.shape[0]), ptr(grad), ptr(step), int(nx),
         int(ny), int(nz), int(x0), int(y0), int(z0), int(sx), int(sy), int(sz), int(from_h),
         stream())


def _reverse_half(q_dst, q_src, eps, psi_in, psi_out, coeff, hist, grad, step, nx, ny, nz,
                  x0, y0, z0, sx, sy, sz, courant, from_h):
    """prepare + gather in one pass (wfd_adjoint_reverse_half), bitwise equal
    to _prepare then _gather without the d streams."""
    t = getattr(grad, "tensor", grad)
    name = ("wfd_adjoint_reverse_half_f64" if t.dtype == torch.float64
            else "wfd_adjoint_reverse_half")
    call(name, ptr(q_dst), ptr(q_src), ptr(eps), ptr(psi_in), ptr(psi_out), ptr(coeff),
         int(coeff.shape[1]), ptr(hist), int(hist.shape[0]), ptr(grad), ptr(step), int(nx),
         int(ny), int(nz), int(x0), int(y0), int(z0), int(sx), int(sy), int(sz),
         float(courant), int(from_h), stream())


def _inject(qe, qh, source, weights, step, nx, ny, nz, x0, y0, z0, sy, sz, old_h_only,
            dim):
    call("wfd_adjoint_inject", ptr(qe), ptr(qh), ptr(source), ptr(weights), ptr(step),
         int(nx), int(ny), int(nz), int(x0), int(y0), int(z0), int(sy), int(sz), int(dim),
         int(old_h_only), int(len(weights)), stream())


def _inject_many(sources, qe, qh, step, nx, ny, nz, old_h_only):
    """Every adjoint source plane of one inject in one launch: sources is a
    ctypes array of InjectSource."""
    call("wfd_adjoint_inject_many", ctypes.cast(sources, ctypes.c_void_p), len(sources),
         ptr(qe), ptr(qh), ptr(step), int(nx), int(ny), int(nz), int(old_h_only), stream())


def _retreat(step):
    call("wfd_advance", ptr(step), -1, stream())


def _write_design(plane, eps, nx, ny, nz, x0, y0, z0, sx, sy, sz):
    call("wfd_write_design", ptr(plane), ptr(eps), int(nx), int(ny), int(nz), int(x0),
         int(y0), int(z0), int(sx), int(sy), int(sz), stream())


def _n(inputs, first):
    return int(inputs[first]) * int(inputs[first + 1]) * int(inputs[first + 2])


record_curl = Kernel("wfd_record_curl", _record_curl,
                     dim=lambda i: int(i[9]) * int(i[10]) * int(i[11]))
prepare = Kernel("wfd_adjoint_prepare", _prepare, dim=lambda i: _n(i, 5))
gather = Kernel("wfd_adjoint_gather", _gather, dim=lambda i: _n(i, 5))
inject = Kernel("wfd_adjoint_inject", _inject, takes_dim=True)
retreat = Kernel("wfd_advance", _retreat, dim=lambda i: 1)
write_design = Kernel("wfd_write_design", _write_design,
                      dim=lambda i: int(i[8]) * int(i[9]) * int(i[10]))


# ---------------------------------------------------------------- driver
class Adjoint(Solver):
    """adjoint.py:231-469 on libwfdtd.

    ``box`` (x0, y0, z0, sx, sy, sz) and the design check default to the
    reference's 20 nm smooth bend; pass ``box`` explicitly (and
    ``require_reference_bend=False``) for other bundles.  The objective is
    the modal transmission T = P(mode_out) / P(mode_in) of the two modal
    detectors, ModeOverlapDetector (the reference's) or multi-wavelength
    DFTPlaneDetector planes, where ``objective_wavelength`` picks the
    wavelength (optional when the planes hold one).
    """

    def __init__(self, case, checkpoint_every=None, grad_dtype=np.float64, box=None,
                 require_reference_bend=True, objective_wavelength=None, **solver_kw):
        solver_kw.setdefault("resident", False)
        super().__init__(case, **solver_kw)
        self.objective_wavelength = objective_wavelength
        self._inject_weights = {}
        if require_reference_bend:
            if (self.shape != (200, 200, 75) or self.meta["settings"]["spacing_nm"] != 20
                    or self.meta["settings"]["parameterization"] != "smooth"):
                raise ValueError("Adjoint driver currently supports the 20 nm smooth bend only")
            settings = self.meta["settings"]
            required = {"time_fs": 200, "pml_cells": 10, "domain_um": 4, "voxel_nm": 20,
                        "scene": "bend"}
            if any(settings.get(key) != value for key, value in required.items()) or \
                    settings.get("pad_domain", False):
                raise ValueError("Adjoint driver requires the unpadded 200 fs reference geometry")
        self.box = tuple(box) if box is not None else (100, 20, 30, 80, 80, 20)
        x0, y0, z0, sx, sy, sz = self.box
        if min(self.box) < 0 or x0 + sx > self.shape[0] or y0 + sy > self.shape[1] or \
                z0 + sz > self.shape[2]:
            raise ValueError("Design box outside the grid")
        self.design_n = sx * sy * sz
        if checkpoint_every is not None and checkpoint_every < 1:
            raise ValueError("checkpoint_every must be positive")
        self.checkpoint_every = checkpoint_every
        K = checkpoint_every or self.steps
        self.history = Array.empty((min(K, self.steps), 3, self.design_n), device=self.device)
        self.qe = Array.zeros((3, self.n), device=self.device)
        self.qh = Array.zeros((3, self.n), device=self.device)
        # adjoint CPML psi: each reverse half-step reads one buffer and writes
        # the other (neighbours still read the old values); _flip says which
        # holds the state, and every captured chunk ends with it in the first
        self.psie = Array.zeros(self.psi_e.shape, device=self.device)
        self.psih = Array.zeros(self.psi_h.shape, device=self.device)
        self.psie2 = Array.zeros(self.psi_e.shape, device=self.device)
        self.psih2 = Array.zeros(self.psi_h.shape, device=self.device)
        self._flip = 0
        self.grad = Array.zeros(self.design_n, np.dtype(grad_dtype), self.device)
        self.reverse_step = Array.zeros(1, np.int32, self.device)
        # captured reverse steps end with one launch for inject(1), the retreat
        # and the next step's inject(0) (wfd_adjoint_inject_turn); False keeps
        # the reference's five launches per step (same bits)
        self.turned_inject = True
        self._turn_ticket = Array.zeros(1, np.int32, self.device)
        self.adjoint_sources = []
        self.reverse_graphs = {}
        self.checkpoints = None
        if checkpoint_every is not None:
            count = (self.steps + K - 1) // K
            shapes = [(3, self.n), (3, self.n), self.psi_e.shape, self.psi_h.shape]
            self.checkpoints = [[Array.empty(s, device=self.device) for s in shapes]
                                for _ in range(count)]

    # --------------------------------------------------------- forward
    def _state(self):
        return [self.E, self.H, self.psi_e, self.psi_h]

    def history_target(self):
        """(history, slots, design box): where a two-step pass writes the curl
        of its steps."""
        return (self.history, self.history.shape[0], self.box)

    def record_curl(self):
        """The curl of the current step into the history (before a fused step)."""
        x0, y0, z0, sx, sy, sz = self.box
        call("wfd_record_curl", ptr(self.current_h()), ptr(self.history),
             int(self.history.shape[0]),
             ptr(self.step), *self.shape, x0, y0, z0, sx, sy, sz, float(self.courant),
             stream())

    def launch_step(self, record=True, curl=True):
        """adjoint.py:271-327: record_curl reads H before this step's H
        half-step, i.e. before the fused E+H step as well."""
        if curl and self._schedule and self._schedule[0] in ("p2", "skip"):
            # a two-step pass records the curl of both its steps itself
            self._curl = self.history_target()
        elif curl:
            self.record_curl()
        self._step_fields(stage=record)
        self._curl = None
        if record:
            self._record()
        self._finish_step()

    def compile(self):
        start = time.perf_counter()
        self.launch_step()
        torch.cuda.synchronize(self.device)
        self.tune()
        self.capture_all()
        return time.perf_counter() - start

    def capture_all(self):
        """The forward, quiet and recompute graphs with the current stepping
        choices (compile's tail)."""
        self.reset()
        self._destroy_graphs()
        full = self.checkpoint_every is None
        lengths = {self.graph_steps, self.steps % self.graph_steps} - {0}
        if not full:
            K = self.checkpoint_every
            lengths |= {K, self.steps % K} - {0}
        for length in lengths:
            # forward with monitors; with checkpoints the curl history is only
            # recorded during recomputation (two-step passes record it themselves)
            self.graphs[length] = self._capture_fn(
                length, lambda: self.launch_step(record=True, curl=full))
            self._quiet_graphs[length] = self._capture_fn(
                length, lambda: self.launch_step(record=False, curl=full))
        self._recompute_graphs = {}
        if not full:
            for length in {self.checkpoint_every, self.steps % self.checkpoint_every} - {0}:
                self._recompute_graphs[length] = self._capture_fn(
                    length, lambda: self.launch_step(record=False, curl=True))
        self.reset()
        self._compiled = True

    def _replay_graphs(self, t0, t1, graphs_by_len, quiet_by_len, active):
        """Replay graphs covering steps [t0, t1) in chunks of the largest length."""
        s = stream()
        t = t0
        lengths = sorted(graphs_by_len, reverse=True)
        while t < t1:
            length = next(L for L in lengths if L <= t1 - t)
            quiet = quiet_by_len.get(length) if quiet_by_len else None
            g = quiet if (quiet is not None and not active[t:t + length].any()) \
                else graphs_by_len[length]
            call("wfd_graph_launch", g, s)
            t += length

    def run(self):
        if self.checkpoint_every is None:
            return super().run()
        K = self.checkpoint_every
        active = self._windows()
        for c, t0 in enumerate(range(0, self.steps, K)):
            for dst, src in zip(self.checkpoints[c], self._state()):
                dst.tensor.copy_(src.tensor)
            t1 = min(t0 + K, self.steps)
            graphs = {L: g for L, g in self.graphs.items() if L in (K, self.steps % K)}
            self._replay_graphs(t0, t1, graphs, self._quiet_graphs, active)
        torch.cuda.synchronize(self.device)

    # --------------------------------------------------------- reverse
    def objective_index(self, m):
        """Wavelength index of a DFT plane the objective reads."""
        if self.objective_wavelength is None:
            if m.nf != 1:
                raise ValueError("objective_wavelength is required with multi-wavelength planes")
            return 0
        hit = np.flatnonzero(np.isclose(m.wavelengths, self.objective_wavelength, rtol=1e-9,
                                        atol=0))
        if not len(hit):
            raise ValueError(f"{m.name} records no wavelength {self.objective_wavelength}")
        return int(hit[0])

    def _modal_terms(self, m):
        """(phasor (6, *plane), mode E, mode H, area, inject weights) of the
        objective's frequency."""
        if hasattr(m, "nf"):  # DFT plane: alpha carries no area (detectors.DFTMonitor)
            f = self.objective_index(m)
            key = (m.name, f)
            if key not in self._inject_weights:
                w = m.weights.numpy()[:, f, :]
                self._inject_weights[key] = Array.from_numpy(w, np.float32, self.device)
            return (m.phasors()[f], self.data[m.name + "/_mode_E"][f],
                    self.data[m.name + "/_mode_H"][f], np.ones(m.shape),
                    self._inject_weights[key])
        return (m.phasors(), self.data[m.name + "/_mode_E"][0], self.data[m.name + "/_mode_H"][0],
                m.area_numpy, m.weights)

    def transmission(self):
        """T = P(mode_out) / P(mode_in) at the objective's wavelength."""
        p = {}
        for m in self.monitors:
            if m.name in ("mode_in", "mode_out"):
                ph, em, hm, area, _ = self._modal_terms(m)
                beta = np.sum(self._overlap(m, em, hm, area) * ph)
                p[m.name] = float(abs(beta) ** 2)
        return p["mode_out"] / p["mode_in"]

    def transmission_spectrum(self):
        """T per wavelength of DFT modal planes mode_in / mode_out."""
        from .detectors import dft_transmission

        return dft_transmission(self)

    @staticmethod
    def _overlap(m, em, hm, area):
        """coeff with sum(coeff * phasor) = conj(alpha) (adjoint.py:343-352)."""
        a, b = (m.axis + 1) % 3, (m.axis + 2) % 3
        qe = np.zeros_like(em)
        qh = np.zeros_like(hm)
        qe[a] = hm[b].conj() / 4
        qe[b] = -hm[a].conj() / 4
        qh[b] = em[a].conj() / 4
        qh[a] = -em[b].conj() / 4
        return np.concatenate((qe, qh)) * area[None]

    def prepare_sources(self):
        """adjoint.py:329-366 — host cotangents, refreshed in place."""
        previous = self.adjoint_sources
        monitors = {m.name: m for m in self.monitors}
        terms = {name: self._modal_terms(monitors[name]) for name in ("mode_in", "mode_out")}
        powers = {}
        for name, (ph, em, hm, area, _) in terms.items():
            beta = np.sum(self._overlap(monitors[name], em, hm, area) * ph)
            powers[name] = float(abs(beta) ** 2)
        pin, pout = powers["mode_in"], powers["mode_out"]
        if not np.isfinite(pin) or pin <= 0 or not np.isfinite(pout):
            raise ValueError("Finite modal powers and positive incident power required")
        self.adjoint_sources = []
        for name, factor in [("mode_in", pout / pin ** 2), ("mode_out", -1 / pin)]:
            m = monitors[name]
            ph, em, hm, area, _ = terms[name]
            coeff = self._overlap(m, em, hm, area)
            beta = np.sum(coeff * ph)
            source = (2 * factor * beta.conjugate() * coeff).reshape(6, -1).astype(np.complex64)
            host = np.stack((source.real, source.imag), axis=-1)
            if previous:
                old_monitor, v = previous[len(self.adjoint_sources)]
                if old_monitor.name != m.name:
                    raise ValueError("Changing detector layout requires new graphs")
                v.assign(host)
            else:
                v = Array.from_numpy(host, np.float32, self.device)
            self.adjoint_sources.append((m, v))

    def inject(self, old):
        """adjoint.py:181-223: add every adjoint source plane's terms to qe, qh
        (one launch for up to WFD_MAX_INJECT planes)."""
        if not self.adjoint_sources:
            return
        if len(self.adjoint_sources) > WFD_MAX_INJECT:
            for m, v in self.adjoint_sources:
                lo = [a for a, b in m.obj["grid_slice"]]
                _inject(self.qe, self.qh, v, self._inject_weights_of(m), self.reverse_step,
                        *self.shape, *lo, *m.shape[1:], old, dim=m.n)
            return
        _inject_many(self._inject_table(), self.qe, self.qh, self.reverse_step, *self.shape,
                     old)

    def _inject_table(self):
        rows = []
        for m, v in self.adjoint_sources:
            lo = [a for a, b in m.obj["grid_slice"]]
            weights = self._inject_weights_of(m)
            rows.append((ptr(v).value, ptr(weights).value, *lo, *m.shape[1:], m.n,
                         len(weights)))
        rows = tuple(rows)
        if getattr(self, "_inject_rows", None) != rows:
            self._inject_rows = rows
            self._inject_sources = (InjectSource * len(rows))(*[InjectSource(*r)
                                                                 for r in rows])
        return self._inject_sources

    def _inject_weights_of(self, m):
        if hasattr(m, "nf"):
            return self._inject_weights[(m.name, self.objective_index(m))]
        return m.weights

    def reverse_one(self):
        """adjoint.py:388-414: inject, prepare H + gather qe, prepare E + gather
        qh (each pair one pass), inject the H_old terms, retreat."""
        self.inject(0)
        if self._flip:
            e_in, e_out, h_in, h_out = self.psie2, self.psie, self.psih2, self.psih
        else:
            e_in, e_out, h_in, h_out = self.psie, self.psie2, self.psih, self.psih2
        _reverse_half(self.qe, self.qh, self.eps, h_in, h_out, self.coeffs["H"], self.history,
                      self.grad, self.reverse_step, *self.shape, *self.box, self.courant, 1)
        _reverse_half(self.qh, self.qe, self.eps, e_in, e_out, self.coeffs["E"], self.history,
                      self.grad, self.reverse_step, *self.shape, *self.box, self.courant, 0)
        self._flip ^= 1
        self.inject(1)
        _retreat(self.reverse_step)

    def _turns(self):
        """Whether the captured reverse steps use wfd_adjoint_inject_turn."""
        return self.turned_inject and len(self.adjoint_sources) <= WFD_MAX_INJECT

    def _reverse_one_turned(self):
        """reverse_one with this step's inject(0) already applied (by the
        previous turn, or by backward() before the first step), ending with
        the turn: inject(1) of this step, retreat, inject(0) of the next."""
        if self._flip:
            e_in, e_out, h_in, h_out = self.psie2, self.psie, self.psih2, self.psih
        else:
            e_in, e_out, h_in, h_out = self.psie, self.psie2, self.psih, self.psih2
        _reverse_half(self.qe, self.qh, self.eps, h_in, h_out, self.coeffs["H"], self.history,
                      self.grad, self.reverse_step, *self.shape, *self.box, self.courant, 1)
        _reverse_half(self.qh, self.qe, self.eps, e_in, e_out, self.coeffs["E"], self.history,
                      self.grad, self.reverse_step, *self.shape, *self.box, self.courant, 0)
        self._flip ^= 1
        table = (ctypes.cast(self._inject_table(), ctypes.c_void_p) if self.adjoint_sources
                 else None)
        call("wfd_adjoint_inject_turn", table, len(self.adjoint_sources), ptr(self.qe),
             ptr(self.qh), ptr(self.reverse_step), ptr(self._turn_ticket), *self.shape,
             stream())

    def _reverse_chunk(self, length):
        """`length` reverse steps leaving the psi state in the first buffers."""
        for _ in range(length):
            if self._turns():
                self._reverse_one_turned()
            else:
                self.reverse_one()
        if self._flip:
            self.psie.tensor.copy_(self.psie2.tensor)
            self.psih.tensor.copy_(self.psih2.tensor)
            self._flip = 0

    def reset_reverse(self):
        for a in (self.qe, self.qh, self.psie, self.psih, self.psie2, self.psih2, self.grad):
            a.zero_()
        self._flip = 0
        self.reverse_step.assign(np.array([self.steps - 1], np.int32))
        torch.cuda.synchronize(self.device)

    def compile_reverse(self):
        if self.cpml_storage != "full":
            raise ValueError("The adjoint needs the full CPML storage")
        if not self.adjoint_sources:
            raise ValueError("prepare_sources() must run before compile_reverse(): the "
                             "reverse graphs capture the cotangent buffers")
        self.reset_reverse()
        self.reverse_one()
        if self._turns():
            self._reverse_one_turned()
        torch.cuda.synchronize(self.device)
        self.reset_reverse()
        for length in self._reverse_lengths():
            self.reverse_graphs[length] = self._capture_fn(
                1, lambda length=length: self._reverse_chunk(length), stepping=False)
        self.reset_reverse()

    def _reverse_lengths(self):
        if self.checkpoint_every is None:
            return {self.graph_steps, self.steps % self.graph_steps} - {0}
        return {self.checkpoint_every, self.steps % self.checkpoint_every} - {0}

    def backward(self):
        self.launch_backward()
        torch.cuda.synchronize(self.device)

    def launch_backward(self):
        """Enqueue the whole reverse on the current stream (no wait): with
        checkpoints, each segment's state restore, recompute and reverse."""
        s = stream()
        if self._turns():
            self.inject(0)  # the first step's; every turn applies the next one's
        if self.checkpoint_every is None:
            for _ in range(self.steps // self.graph_steps):
                call("wfd_graph_launch", self.reverse_graphs[self.graph_steps], s)
            if self.steps % self.graph_steps:
                call("wfd_graph_launch", self.reverse_graphs[self.steps % self.graph_steps], s)
            return
        K = self.checkpoint_every
        starts = list(range(0, self.steps, K))
        for c in reversed(range(len(starts))):
            t0 = starts[c]
            length = min(K, self.steps - t0)
            for dst, src in zip(self._state(), self.checkpoints[c]):
                dst.tensor.copy_(src.tensor)
            self.counters.tensor[0] = t0
            call("wfd_graph_launch", self._recompute_graphs[length], s)
            call("wfd_graph_launch", self.reverse_graphs[length], s)

    def set_design(self, parameters):
        from .material import Material

        material = Material(np.asarray(parameters).reshape(80, 80), self.meta["beta"],
                            device=self.device)
        x0, y0, z0, sx, sy, sz = self.box
        _write_design(material.inv_eps, self.eps, *self.shape, *self.box)
        return material

    def differentiate(self, parameters):
        """adjoint.py:453-469 — (-T, dT/dp) for a new design."""
        from .detectors import measurements

        material = self.set_design(parameters)
        if not self.graphs:
            self.compile()
        self.reset()
        self.run()
        self.prepare_sources()
        if not self.reverse_graphs:
            self.compile_reverse()
        self.reset_reverse()
        self.backward()
        x0, y0, z0, sx, sy, sz = self.box
        voxel = self.grad.numpy().reshape(sx, sy, sz).sum(axis=2)
        gradient = material.pullback(voxel)
        return -measurements(self)["modal_transmission"], gradient

    def voxel_gradient(self):
        x0, y0, z0, sx, sy, sz = self.box
        return self.grad.numpy().reshape(sx, sy, sz).sum(axis=2)
