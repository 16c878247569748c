"""Launch-shape autotuner for the fused Yee step.

The only free launch parameter of wfd_yee_step is the x chunk a thread column
marches (1, 2, 4 or 8 planes, compile-time unrolled).  ``tune`` times every
candidate on the solver's own grid and buffers with HIP events and returns
the fastest; results are cached in memory and in a JSON file keyed by grid
shape, CPML rows and device name, so a shape is tuned once per machine.
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
