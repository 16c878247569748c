"""Many independent grids stepped together (no reference counterpart).

``Batch`` holds one ``Solver`` per design (any shapes, the same number of
steps) and advances all of them with ``wfd_batch_step``: one launch per
half-step and x-chunk class covers every design's blocks, each design keeps
its own fields, CPML, source, counters and staged detectors, and its results
are bitwise those of stepping it alone.  ``compile`` tunes every design's x
chunk (autotune.py, cached by shape), then times the batched step with those
and with one common chunk for all designs and keeps the fastest, writes the
device argument tables and
captures one HIP graph per chunk of ``graph_steps`` steps that steps every
design and replays every design's staged detectors; ``run`` replays them.
Where every design can take the single pass, ``compile`` also times the
two-step pass, as one batched launch (``wfd_batch_pass``: two steps of every
design per launch) and as one ``wfd_yee_pass`` per design in turn (each
sized for the whole GPU), each design alternating between its own two buffer
sets, and uses the fastest of the three; chunks end back in each design's
first set.  ``single_pass``: "auto", True (a pass form even if slower),
"batched" or "each" (that form), False (the fused step only).  Adjoint
designs without checkpoints (``checkpoint_every`` None) record their whole
curl history during the batched forward: the two-step pass writes it
itself, and a fused step follows one record_curl launch per such design.
"""

import ctypes
import time

import numpy as np
import torch

from ._lib import WFD_MAX_BATCH, Handles, YeePlan, YeeTargets, call, ptr, release, stream
from .array import Array


class Batch:
    def __init__(self, solvers, graph_steps=64, single_pass="auto"):
        self.solvers = list(solvers)
        self.pass_mode = single_pass
        if not 1 <= len(self.solvers) <= WFD_MAX_BATCH:
            raise ValueError(f"A batch holds 1..{WFD_MAX_BATCH} solvers")
        steps = {s.steps for s in self.solvers}
        if len(steps) != 1:
            raise ValueError("Every design of a batch must run the same number of steps")
        if any(s.cpml_storage != "full" for s in self.solvers):
            raise ValueError("Batched stepping needs the full CPML storage")
        if any(len(s.staged()) != len(s.monitors) for s in self.solvers):
            raise ValueError("Batched stepping records staged detectors only")
        if graph_steps < 1 or any(s.graph_steps != graph_steps for s in self.solvers):
            raise ValueError("graph_steps must be positive and equal to every solver's")
        self.steps = steps.pop()
        self.graph_steps = graph_steps
        self.device = self.solvers[0].device
        self.graphs, self.quiet_graphs = {}, {}
        self._handles = Handles()
        self._tables = {}
        self._pass_tables = {}
        self.use_pass = False
        self.pass_each = False  # passes one design at a time instead of one batched launch
        self.pass_us = None
        self.cells = sum(s.n for s in self.solvers)

    # ------------------------------------------------------------ tables
    def _table(self, record):
        plans = (YeePlan * len(self.solvers))(*[s._plan(stage=record) for s in self.solvers])
        nbytes = int(call_bytes(len(self.solvers)))
        ws = Array.zeros((nbytes + 3) // 4, np.int32, self.device)
        call("wfd_batch_prepare", plans, len(self.solvers), ptr(ws), nbytes, stream())
        return plans, ws

    def _full_history(self):
        """The designs whose forward records the whole curl history (adjoints
        without checkpoints): every batched step records it for them."""
        return [s for s in self.solvers
                if hasattr(s, "history_target") and s.checkpoint_every is None]

    def _pass_table(self, record, direction):
        """Plans and targets of the batched two-step pass from each design's
        first buffer set to its second (direction 0) or back (1); the pass
        writes the curl history of designs that keep it whole."""
        full = set(map(id, self._full_history()))
        plans, targets = [], []
        for solver in self.solvers:
            plan = solver._plan(stage=record)
            main = {"E": solver.E, "H": solver.H, "psi_e": solver.psi_e, "psi_h": solver.psi_h}
            src, dst = (main, solver._alt) if direction == 0 else (solver._alt, main)
            plan.e, plan.h = ptr(src["E"]), ptr(src["H"])
            plan.psi_e, plan.psi_h = ptr(src["psi_e"]), ptr(src["psi_h"])
            out = YeeTargets()
            out.e, out.h = ptr(dst["E"]), ptr(dst["H"])
            out.psi_e, out.psi_h = ptr(dst["psi_e"]), ptr(dst["psi_h"])
            out.ticket = ptr(solver._ticket)
            out.steps, out.slots_per_thread, out.threads = 2, 1, 1024
            if id(solver) in full:
                hist, slots, box = solver.history_target()
                out.hist, out.hist_steps = ptr(hist), int(slots)
                for q, v in enumerate(box):
                    out.hist_box[q] = int(v)
            plans.append(plan)
            targets.append(out)
        n = len(self.solvers)
        plans = (YeePlan * n)(*plans)
        targets = (YeeTargets * n)(*targets)
        from ._lib import lib

        nbytes = int(lib().wfd_batch_pass_workspace_bytes(n))
        ws = Array.zeros((nbytes + 3) // 4, np.int32, self.device)
        call("wfd_batch_pass_prepare", plans, targets, n, ptr(ws), nbytes, stream())
        return plans, targets, ws

    def _pass_launch(self, record, direction, each=None):
        """Two steps of every design: one batched launch, or (``each``) one
        wfd_yee_pass per design in turn, each sized for the whole GPU."""
        plans, targets, ws = self._pass_tables[(record, direction)]
        if self.pass_each if each is None else each:
            for d in range(len(self.solvers)):
                call("wfd_yee_pass", ctypes.byref(plans[d]), ctypes.byref(targets[d]), stream())
        else:
            call("wfd_batch_pass", plans, targets, len(self.solvers), ptr(ws), stream())

    def _fused_step(self, plans, ws, full):
        """One batched fused step; the designs in ``full`` record their curl
        first, as their own forward does."""
        for solver in full:
            solver.record_curl()
        call("wfd_batch_step", plans, len(self.solvers), ptr(ws), stream())

    def pass_applies(self):
        return all(s.single_pass_applies() for s in self.solvers)

    def tune_pass(self, steps=8):
        """Microseconds per step of the two-step pass (one batched launch, and
        one launch per design) and of the batched fused step (quiet tables,
        HIP events); keeps the faster pass form.  Returns (pass, fused) with
        the faster pass form's time."""
        for s in self.solvers:
            single = s.single_pass
            s.use_single_pass(0, 1, 2, 1024)  # allocates the second buffer set and ticket
            s.single_pass = single
        self._pass_tables = {(r, d): self._pass_table(r, d) for r in (True, False) for d in (0, 1)}
        plans, ws = self._table(False)
        full = self._full_history()
        timings = []
        # (run, steps it advances): a pass launch advances two steps, so a
        # pair there and back advances four.
        for run, advance in (
                (lambda: [self._pass_launch(False, 0, False), self._pass_launch(False, 1, False)], 4),
                (lambda: [self._pass_launch(False, 0, True), self._pass_launch(False, 1, True)], 4),
                (lambda: [self._fused_step(plans, ws, full) for _ in range(2)], 2)):
            run()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            repeats = max(1, steps // advance)
            for _ in range(repeats):
                run()
            b.record()
            torch.cuda.synchronize(self.device)
            timings.append(a.elapsed_time(b) * 1e3 / (advance * repeats))
        self.reset()
        self.pass_each = timings[1] < timings[0]
        self.pass_us = {"batched pass": round(timings[0], 1),
                        "pass per design": round(timings[1], 1), "fused": round(timings[2], 1)}
        return min(timings[0], timings[1]), timings[2]

    def step(self, record=True):
        """One step of every design (eager); staged detectors replay each step."""
        plans, ws = self._tables[record]
        self._fused_step(plans, ws, self._full_history())
        if record:
            for s in self.solvers:
                s._replay_staged(1)

    def _capture(self, length, record):
        plans, ws = self._tables[record]
        full = self._full_history()
        torch.cuda.synchronize(self.device)
        s = self._handles.begin()
        ext = torch.cuda.ExternalStream(s.value, device=self.device)
        try:
            with torch.cuda.stream(ext):
                done = 0
                while done < length:
                    m = min(self.graph_steps, length - done)
                    # two-step passes in pairs (back in the first sets), the
                    # rest as fused steps
                    passes = 2 * (m // 4) if self.use_pass else 0
                    for k in range(passes):
                        self._pass_launch(record, k % 2)
                    for _ in range(m - 2 * passes):
                        self._fused_step(plans, ws, full)
                    if record:
                        for solver in self.solvers:
                            solver._replay_staged(m)
                    done += m
        finally:
            graph = self._handles.end()
        return graph

    # ------------------------------------------------------------ driver
    def tune_x_chunk(self, candidates=(1, 2, 4, 8), steps=6):
        """Time the whole batched step (quiet tables, HIP events) with every
        design at each x chunk and keep the fastest common chunk if it beats
        the designs' own tuned chunks.  Returns {label: microseconds}."""
        own = [s.x_chunk for s in self.solvers]
        options = [("own", own)] + [(str(c), [c] * len(self.solvers)) for c in candidates]
        timings = {}
        for label, chunks in options:
            for s, c in zip(self.solvers, chunks):
                s.x_chunk = c
            plans, ws = self._table(False)
            call("wfd_batch_step", plans, len(self.solvers), ptr(ws), stream())
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(steps):
                call("wfd_batch_step", plans, len(self.solvers), ptr(ws), stream())
            b.record()
            torch.cuda.synchronize(self.device)
            timings[label] = a.elapsed_time(b) * 1e3 / steps
        best = min(options, key=lambda o: timings[o[0]])
        for s, c in zip(self.solvers, best[1]):
            s.x_chunk = c
        for s in self.solvers:
            s.reset()
        return timings

    def compile(self):
        start = time.perf_counter()
        for s in self.solvers:
            s.resident_mode = False
            s.launch_step()
            torch.cuda.synchronize(self.device)
            s.tune()
            s.reset()
        self.x_chunk_us = self.tune_x_chunk()
        if self.pass_mode is not False and self.pass_applies():
            pass_us, fused_us = self.tune_pass()
            self.use_pass = self.pass_mode != "auto" or pass_us < fused_us
            if self.pass_mode in ("batched", "each"):
                self.pass_each = self.pass_mode == "each"
        self._destroy_graphs()
        self._tables = {True: self._table(True), False: self._table(False)}
        for length in {self.graph_steps, self.steps % self.graph_steps} - {0}:
            self.graphs[length] = self._capture(length, True)
            self.quiet_graphs[length] = self._capture(length, False)
        self.reset()
        return time.perf_counter() - start

    def reset(self):
        for s in self.solvers:
            s.reset()

    def run(self):
        """Every design's full trajectory: one graph launch per chunk."""
        if not self.graphs:
            raise RuntimeError("compile() must run before run()")
        active = np.zeros(self.steps, bool)
        for s in self.solvers:
            active |= s._windows()
        t = 0
        for length in [self.graph_steps] * (self.steps // self.graph_steps) + (
                [self.steps % self.graph_steps] if self.steps % self.graph_steps else []):
            g = self.graphs[length] if active[t:t + length].any() else self.quiet_graphs[length]
            call("wfd_graph_launch", g, stream())
            t += length
        torch.cuda.synchronize(self.device)

    def run_steps(self, count):
        """Advance every design exactly `count` steps from its current step
        (recorded graphs of at most graph_steps; missing lengths are captured
        on first use) and synchronize."""
        if not self.graphs:
            raise RuntimeError("compile() must run before run_steps()")
        left = int(count)
        while left > 0:
            length = min(self.graph_steps, left)
            if length not in self.graphs:
                self.graphs[length] = self._capture(length, True)
            call("wfd_graph_launch", self.graphs[length], stream())
            left -= length
        torch.cuda.synchronize(self.device)

    def _destroy_graphs(self):
        for table in (self.graphs, self.quiet_graphs):
            for g in table.values():
                self._handles.destroy(g)
            table.clear()

    def close(self):
        """Destroy the batch's HIP graphs and capture stream (idempotent)."""
        self.graphs.clear()
        self.quiet_graphs.clear()
        self._handles.close()

    def __del__(self, _release=release):
        # (bound at definition: module globals may already be None when an
        # object is collected during interpreter shutdown)
        _release(getattr(self, "_handles", None))


def call_bytes(count):
    from ._lib import lib

    nbytes = lib().wfd_batch_workspace_bytes(int(count))
    if nbytes < 0:
        raise ValueError(f"A batch holds 1..{WFD_MAX_BATCH} plans")
    return nbytes
