"""Launch-shape autotuner for the Yee time step.

The free launch parameter of wfd_yee_step is the x chunk a thread column
marches (1, 2, 4 or 8 planes, compile-time unrolled).  ``tune`` times every
candidate on the solver's own grid and buffers with HIP events and returns
the fastest; ``tune_resident`` times the on-chip-resident kernel
(wfd_resident_run) against the tuned fused step.  Results are cached in
memory and in a JSON file keyed by grid shape, CPML rows and device name, so
a shape is tuned once per machine.
"""

import ctypes
import json
import os
from pathlib import Path

import numpy as np
import torch

from ._lib import call, stream

CANDIDATES = (1, 2, 4, 8)
_memory = {}


def cache_path():
    root = os.environ.get("WFDTD_CACHE") or os.path.join(
        os.environ.get("XDG_CACHE_HOME") or os.path.expanduser("~/.cache"), "wfdtd")
    return Path(root) / "autotune.json"


def _key(solver):
    name = torch.cuda.get_device_name(solver.device)
    return f"{name}|{solver.shape}|{solver.coeffs['E'].shape[1]}|yee_step"


def _load():
    try:
        return json.loads(cache_path().read_text())
    except (OSError, ValueError):
        return {}


def _store(table):
    try:
        path = cache_path()
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(".tmp")
        tmp.write_text(json.dumps(table, indent=1))
        os.replace(tmp, path)
    except OSError:
        pass


def measure(solver, x_chunk, repeats=40):
    """Median microseconds of one fused step with the given x chunk."""
    plan = solver._plan()
    plan.x_chunk = int(x_chunk)
    for _ in range(5):
        call("wfd_yee_step", ctypes.byref(plan), stream())
    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(repeats)]
    for a, b in events:
        a.record()
        call("wfd_yee_step", ctypes.byref(plan), stream())
        b.record()
    torch.cuda.synchronize(solver.device)
    return float(np.median([a.elapsed_time(b) for a, b in events]) * 1e3)


def tune(solver, use_cache=True):
    """Fastest x chunk for the solver's grid (runs real steps: reset after)."""
    key = _key(solver)
    if use_cache:
        if key in _memory:
            return _memory[key]
        table = _load()
        if key in table:
            _memory[key] = int(table[key]["x_chunk"])
            return _memory[key]
    timings = {xc: measure(solver, xc) for xc in CANDIDATES}
    best = min(timings, key=timings.get)
    _memory[key] = best
    table = _load()
    table[key] = {"x_chunk": best, "us": {str(k): round(v, 3) for k, v in timings.items()}}
    _store(table)
    return best


def measure_resident(solver, layout, steps=64, repeats=5):
    """Median microseconds per step of resident launches of `steps` steps
    (monitors not recorded); leaves the solver in streaming mode."""
    solver.use_resident(layout)
    try:
        solver.resident_steps(steps, record=False)
        events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                  for _ in range(repeats)]
        for a, b in events:
            a.record()
            solver.resident_steps(steps, record=False)
            b.record()
        solver.check()
    finally:
        solver.resident = False
    return float(np.median([a.elapsed_time(b) for a, b in events]) * 1e3 / steps)


def tune_resident(solver, layout, use_cache=True):
    """True when the resident kernel beats the tuned fused step on this grid."""
    key = _key(solver) + f"|resident|{layout.sx}x{layout.sy}x{layout.sz}"
    if use_cache:
        if key in _memory:
            return _memory[key]
        table = _load()
        if key in table:
            _memory[key] = bool(table[key]["resident"])
            return _memory[key]
    fused = measure(solver, solver.x_chunk or 2)
    resident = measure_resident(solver, layout)
    best = resident < fused
    _memory[key] = best
    table = _load()
    table[key] = {"resident": best, "us": {"fused": round(fused, 3),
                                           "resident": round(resident, 3)}}
    _store(table)
    return best
