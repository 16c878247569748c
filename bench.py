#!/usr/bin/env python3
"""Headline benchmark of the uniform-grid FDTD path on one MI355X.

Workload: the 160x96x80 straight silicon waveguide at 20 nm (500x220 nm Si
core in oxide, 12-cell CPML on all faces, TE0 modal source from the
project's eigenmode solver with a Hann-windowed pulse at 1.55 um, modal +
flux DFT planes at 8 wavelengths 1.50-1.60 um), 4000 time steps.  A bench
"step" is one Yee time step (E and H of every cell, CPML, source and the DFT
detectors); inputs are resident in HBM before the timed region.

    python bench.py [--gpus 1] [--steps 4000] [--warmup 256] [--no-cpu]

Prints one JSON line: Mcell-updates/s, the roofline of the fused Yee step
kernels (algorithmic bytes B = 52 + 32 f_cpml per cell-update, f_cpml =
sum over axes of 2L/N, over the measured kernel time, against 8 TB/s) and
the CPU oracle's rate on a bounded sample of the same grid.
"""

import argparse
import json
import os
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

SHAPE = (160, 96, 80)
PML = 12
SPACING = 20e-9
STEPS = 4000
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_cell(shape, pml):
    f_cpml = sum(2 * pml / n for n in shape)
    return 52.0 + 32.0 * f_cpml, f_cpml


def build_case():
    import __graft_entry__

    pkg = __graft_entry__.package()
    path = Path(tempfile.gettempdir()) / "wfdtd_bench_waveguide_160x96x80"
    if not (path / "manifest.json").exists():
        pkg.cases.waveguide_case(path, shape=SHAPE, spacing=SPACING, steps=STEPS, pml=PML)
    return pkg, path


def kernel_time_per_step(pkg, s, repeats, chunk=64):
    """Average duration of one Yee step from HIP events recorded on the launch
    stream: around each fused E + H kernel pair, or around resident launches
    of `chunk` steps (divided by chunk)."""
    import torch

    if s.resident:
        pairs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                 for _ in range(max(4, repeats // 8))]
        s.resident_steps(chunk, record=False)
        for a, b in pairs:
            a.record()
            s.resident_steps(chunk, record=False)
            b.record()
        s.check()
        times = np.array([a.elapsed_time(b) for a, b in pairs]) * 1e-3 / chunk
        return float(np.median(times)), float(np.mean(times))
    plan = s._plan()
    pairs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
             for _ in range(repeats)]
    for _ in range(8):
        pkg.kernels.yee_step(plan)
    for a, b in pairs:
        a.record()
        pkg.kernels.yee_step(plan)
        b.record()
    torch.cuda.synchronize()
    times = np.array([a.elapsed_time(b) for a, b in pairs]) * 1e-3
    return float(np.median(times)), float(np.mean(times))


def cpu_baseline(s, budget_s=12.0):
    """The CPU oracle (a scalar OpenMP restatement of the reference's
    update_e/update_h) on the same grid and tables for a bounded number of
    steps; detectors are plane-sized and left out of the sample."""
    from oracle import kernels as OK

    nx, ny, nz = s.shape
    n = s.n
    E = np.zeros((3, n), np.float32)
    H = np.zeros((3, n), np.float32)
    hprev = np.zeros((3, 1), np.float32)
    eps = s.eps.numpy()
    psi = [np.zeros((6, n), np.float32) for _ in range(2)]
    coeff = [np.ascontiguousarray(s._coeff_rows[f], np.float32) for f in "EH"]
    src = {f: s.sources[f].numpy() for f in "EH"}
    wav = {f: s.waves[f].numpy() for f in "EH"}
    step = np.zeros(1, np.int32)

    def one():
        OK.update(True, E, H, eps, psi[0], coeff[0], src["H"], wav["H"], step, hprev,
                  s.steps + 1, s.shape, s.source_x, s.courant, 0)
        OK.update(False, H, E, eps, psi[1], coeff[1], src["E"], wav["E"], step, hprev,
                  s.steps + 1, s.shape, s.source_x, s.courant, 0)
        step[0] += 1

    t = time.perf_counter()
    for _ in range(3):
        one()
    probe = (time.perf_counter() - t) / 3
    count = int(max(5, min(2000, budget_s / max(probe, 1e-6))))
    t = time.perf_counter()
    for _ in range(count):
        one()
    elapsed = time.perf_counter() - t
    cores = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    return {"value": n * count / elapsed / 1e6, "unit": "Mcell-updates/s", "cores": cores,
            "kind": "port",
            "sample": f"oracle update_e+update_h, {nx}x{ny}x{nz} grid, steps 3..{3 + count} "
                      f"({elapsed:.1f} s), float32, OpenMP"}


def copy_probe(mib):
    """8 plain device copies of `mib` MiB: reads and writes mib MiB each, the
    known traffic against which FETCH_SIZE / WRITE_SIZE are scaled."""
    import torch

    a = torch.ones(mib * 2 ** 18, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(8):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_mib": mib, "copies": 8}), flush=True)


def main():
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=STEPS)
    p.add_argument("--warmup", type=int, default=256)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle sample")
    p.add_argument("--kernel-repeats", type=int, default=200)
    p.add_argument("--x-chunk", type=int, default=0,
                   help="planes per thread column of the fused step (0: autotune)")
    p.add_argument("--resident", choices=("auto", "on", "off"), default="auto",
                   help="on-chip-resident stepping (auto: the faster of the two)")
    p.add_argument("--copy-mib", type=int, default=0,
                   help="only run 8 device-to-device copies of this many MiB (the reference "
                        "traffic for calibrating HBM counters) and exit")
    args = p.parse_args()
    if args.copy_mib:
        copy_probe(args.copy_mib)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this build "
              "targets a single MI355X; multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.steps > STEPS or args.warmup < 0:
        p.error(f"--steps must be in 1..{STEPS}, --warmup >= 0")
    import torch

    pkg, case = build_case()
    resident = {"auto": "auto", "on": True, "off": False}[args.resident]
    s = pkg.solver.Solver(case, x_chunk=args.x_chunk, resident=resident)
    pkg.detectors.attach(s)
    s.compile()
    if args.warmup:
        s.reset()
        s.run_steps(args.warmup)
    s.reset()
    torch.cuda.synchronize()
    start_ev = torch.cuda.Event(enable_timing=True)
    end_ev = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start_ev.record()
    s.run_steps(args.steps, synchronize=False)
    end_ev.record()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    s.check()
    gpu_elapsed = start_ev.elapsed_time(end_ev) * 1e-3
    cells = s.n
    value = cells * args.steps / elapsed / 1e6
    finite = bool(np.isfinite(s.E.numpy()).all() and np.isfinite(s.H.numpy()).all())

    B, f_cpml = algorithmic_bytes_per_cell(s.shape, PML)
    median_s, mean_s = kernel_time_per_step(pkg, s, args.kernel_repeats)
    achieved = cells * B / mean_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": ("k_resident (steps per launch: 64)" if s.resident
                           else "k_yee_e + k_yee_h (one fused step)"),
                "bytes_per_cell_update": round(B, 2), "f_cpml": round(f_cpml, 4),
                "kernel_us_per_step_mean": round(mean_s * 1e6, 3),
                "kernel_us_per_step_median": round(median_s * 1e6, 3),
                "x_chunk": s.x_chunk, "resident": bool(s.resident)}
    if s.resident:
        L = s.layout
        roofline["resident_layout"] = {"sub_box": [L.sx, L.sy, L.sz],
                                       "workgroups": [L.gx, L.gy, L.gz]}
    line = {
        "metric": "Mcell-updates/s + % HBM roofline (fused Yee+CPML step, 160x96x80 waveguide)",
        "value": round(value, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "straight Si waveguide 160x96x80 @ 20 nm, 12-cell CPML, TE0 "
                               "Hann pulse at 1.55 um, 8-wavelength modal+flux DFT planes",
                   "grid": list(s.shape), "cells": cells, "time_steps": args.steps,
                   "gpu_seconds": round(gpu_elapsed, 6), "finite": finite},
        "roofline": roofline,
        "cpu_baseline": None if args.no_cpu else cpu_baseline(s),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
