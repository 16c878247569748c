"""Launch-shape autotuner for the Yee time step.

The free launch parameter of wfd_yee_step is the x chunk a thread column
marches (1, 2, 4 or 8 planes, compile-time unrolled).  ``tune`` times every
candidate on the solver's own grid and buffers with HIP events and returns
the fastest; ``tune_resident`` times the on-chip-resident kernel
(wfd_resident_run) against the tuned fused step, and ``tune_pass`` the
single-pass step (wfd_yee_pass: x planes per workgroup, slots per thread).  Results are cached in
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


def candidate_layouts(solver, count=5):
    """Resident sub-box candidates: for each (sy, sz) the library accepts, its
    layout with the fewest x planes; the `count` best by (planes, surface),
    and the best two with 2 and 4 times their x planes."""
    nx, ny, nz = solver.shape
    seen, out = set(), []
    zs = sorted({v for v in range(4, nz + 1, 4)} | {nz})
    ys = sorted({v for v in range(2, ny + 1, 2)} | {ny})
    for sz in zs:
        for sy in ys:
            if sy * sz > 1024:
                break
            try:
                L = solver.resident_layout(0, sy, sz)
            except ValueError:
                continue
            key = (L.sx, L.sy, L.sz)
            if key in seen:
                continue
            seen.add(key)
            out.append(L)
    out.sort(key=lambda L: (L.sx, L.sx * L.sy + L.sx * L.sz + L.sy * L.sz, L.blocks))
    out = out[:count]
    # deeper sub-boxes (fewer workgroups, more x planes per column): on small
    # grids they hide the x-face exchange behind the interior planes
    for L in list(out[:2]):
        for factor in (2, 4):
            try:
                M = solver.resident_layout(L.sx * factor, L.sy, L.sz)
            except ValueError:
                continue
            key = (M.sx, M.sy, M.sz)
            if key not in seen:
                seen.add(key)
                out.append(M)
    return out


def tune_resident(solver, use_cache=True, require=False):
    """The fastest resident layout if it beats the tuned fused step on this
    grid (or in any case with ``require``), else None.  Candidates: the
    library's own choice and the best few (sy, sz) factorisations
    (candidate_layouts), each timed on the grid."""
    # (keyed apart from results timed over shallower candidate sets)
    key = _key(solver) + "|resident-layout-deep"
    if use_cache:
        hit = _memory.get(key)
        if hit is None:
            hit = _load().get(key)
        if hit is not None:
            _memory[key] = hit
            if not (hit["resident"] or require):
                return None
            try:
                return solver.resident_layout(*hit["layout"])
            except ValueError:
                pass
    try:
        layouts = [solver.resident_layout()]
    except ValueError:
        return None
    for L in candidate_layouts(solver):
        if (L.sx, L.sy, L.sz) != (layouts[0].sx, layouts[0].sy, layouts[0].sz):
            layouts.append(L)
    fused = measure(solver, solver.x_chunk or 2)
    timings = {f"{L.sx}x{L.sy}x{L.sz}": measure_resident(solver, L) for L in layouts}
    best = min(layouts, key=lambda L: timings[f"{L.sx}x{L.sy}x{L.sz}"])
    best_us = timings[f"{best.sx}x{best.sy}x{best.sz}"]
    use = best_us < fused
    entry = {"resident": use, "layout": [best.sx, best.sy, best.sz],
             "us": {"fused": round(fused, 3), **{k: round(v, 3) for k, v in timings.items()}}}
    _memory[key] = entry
    table = _load()
    table[key] = entry
    _store(table)
    return best if (use or require) else None


# steps per launch -> (x planes per workgroup, slots per thread) candidates
# (0: the library's choice from the kernel's occupancy)
PASS_PLANES = {1: (0, 2, 4, 8), 2: (0, 4, 8, 16)}
# (slots per thread, threads per workgroup) of each steps-per-launch
PASS_SHAPES = {1: ((1, 512), (2, 256), (1, 1024)), 2: ((1, 512), (2, 512), (1, 1024))}


def _pass_fits(nz, steps, slots, threads):
    return (nz if steps == 1 else 3 * nz) < slots * threads


def measure_pass(solver, planes, slots, steps=1, threads=0, repeats=20, table=True):
    """Median microseconds per step of single-pass launches (a pair there
    and back) with the given shape and form; runs real steps (reset after)."""
    solver.use_single_pass(planes, slots, steps, threads, table)
    kinds = ["p2", "skip", "p2", "skip"] if steps == 2 else ["p1", "p1"]
    try:
        def pair():
            solver._schedule = list(kinds)
            for _ in kinds:
                solver._pass_step(False)

        for _ in range(3):
            pair()
        events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                  for _ in range(repeats)]
        for a, b in events:
            a.record()
            pair()
            b.record()
        torch.cuda.synchronize(solver.device)
    finally:
        solver.single_pass = False
        solver._phase, solver._schedule = 0, []
    return float(np.median([a.elapsed_time(b) for a, b in events]) * 1e3 / len(kinds))


def tune_pass(solver, use_cache=True, require=False):
    """(planes, slots per thread, steps per launch, threads, table) of the
    fastest single-pass shape and form if it beats the tuned fused step on
    this grid (or in any case with ``require``), else None.  The shapes are
    timed in the table form where it applies (two-step passes of one slot per
    1024-thread workgroup, launched as a batch of one); the best is then timed
    with its arguments in the kernarg segment too, and the faster form kept."""
    key = _key(solver) + "|single-pass-7"
    if use_cache:
        hit = _memory.get(key)
        if hit is None:
            hit = _load().get(key)
        if hit is not None:
            _memory[key] = hit
            return tuple(hit["shape"]) if (hit["single_pass"] or require) else None
    nz = solver.shape[2]
    shapes = [(p, j, k, t) for k, forms in PASS_SHAPES.items() for j, t in forms
              if _pass_fits(nz, k, j, t) for p in PASS_PLANES[k]]
    if not shapes:
        if require:
            raise ValueError(f"no single-pass shape fits nz = {nz}")
        return None
    fused = measure(solver, solver.x_chunk or 2)
    timings = {"x".join(map(str, sh)): measure_pass(solver, *sh) for sh in shapes}
    best = min(shapes, key=lambda sh: timings["x".join(map(str, sh))])
    best_us = timings["x".join(map(str, best))]
    table = best[2] == 2 and best[1] in (0, 1) and best[3] in (0, 1024)
    if table:
        kernarg = measure_pass(solver, *best, table=False)
        timings["x".join(map(str, best)) + " kernarg"] = kernarg
        table = best_us <= kernarg
        best_us = min(best_us, kernarg)
    best = (*best, table)
    use = best_us < fused
    entry = {"single_pass": use, "shape": list(best),
             "us": {"fused": round(fused, 3), **{k: round(v, 3) for k, v in timings.items()}}}
    _memory[key] = entry
    table = _load()
    table[key] = entry
    _store(table)
    return best if (use or require) else None
