#!/usr/bin/env python3
"""Headline benchmark of the uniform-grid FDTD path on one MI355X.

Workload: the 160x96x80 straight silicon waveguide at 20 nm (500x220 nm Si
core in oxide, 12-cell CPML on all faces, TE0 modal source from the
project's eigenmode solver with a Hann-windowed pulse at 1.55 um, modal +
flux DFT planes at 8 wavelengths 1.50-1.60 um), 4000 time steps.  A bench
"step" is one complete run of that workload -- reset, then 4000 Yee time
steps (E and H of every cell, CPML, source and the DFT detectors) -- as the
reference's benchmark times complete trajectories; inputs are resident in
HBM before the timed region.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--no-cpu]

Prints one JSON line: Mcell-updates/s (cells x 4000 x runs / wall time), the
roofline of the step kernel (algorithmic bytes B = 52 + 32 f_cpml per
cell-update, f_cpml = sum over axes of 2L/N, per launch over the launch's
duration from HIP events on its stream, against 8 TB/s) and the CPU oracle's
rate on a bounded sample of the same grid.
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


# HBM bytes per step measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on this
# workload (scaled by a 256 MiB copy: FETCH x2.000, WRITE x1.000), keyed by
# (shape, resident, steps per launch) -> (bytes per step, the profile summary);
# launches of another length report no traffic.
MEASURED_TRAFFIC = {
    (SHAPE, True, 1000): (44.57e6, "profiles/hbm_traffic_waveguide_160x96x80.md"),
    (SHAPE, False, 1): (130.2e6, "profiles/hbm_traffic_waveguide_160x96x80.md"),
}


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


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or the platform string."""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or platform.machine()


def cpu_baseline(s, budget_s=8.0, single_budget_s=5.0):
    """The CPU oracle (a scalar OpenMP restatement of the reference's
    update_e/update_h) on the same grid and tables for a bounded number of
    steps, on every host core and on one; detectors are plane-sized and left
    out of the sample."""
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

    def rate(budget):
        t = time.perf_counter()
        for _ in range(2):
            one()
        probe = (time.perf_counter() - t) / 2
        count = int(max(3, min(2000, budget / max(probe, 1e-6))))
        first = int(step[0])
        t = time.perf_counter()
        for _ in range(count):
            one()
        elapsed = time.perf_counter() - t
        return n * count / elapsed / 1e6, first, count, elapsed

    cores = len(os.sched_getaffinity(0))
    threads = int(os.environ.get("OMP_NUM_THREADS") or cores)
    previous = OK.set_threads(threads)
    try:
        value, first, count, elapsed = rate(budget_s)
        OK.set_threads(1)
        one_value, one_first, one_count, one_elapsed = rate(single_budget_s)
    finally:
        OK.set_threads(previous)
    return {"value": round(value, 1), "unit": "Mcell-updates/s", "cores": threads,
            "kind": "port", "cpu_model": cpu_model(), "cpus_available": cores,
            "one_thread_value": round(one_value, 2),
            "sample": f"oracle update_e+update_h (float32, OpenMP), {nx}x{ny}x{nz} grid: "
                      f"steps {first}..{first + count} on {threads} threads ({elapsed:.1f} s), "
                      f"then steps {one_first}..{one_first + one_count} on 1 thread "
                      f"({one_elapsed:.1f} s)"}


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


def copy_ceiling(mib=1024, repeats=10):
    """Measured device-copy bandwidth in GB/s (read + write bytes / time) of
    `mib` MiB buffers, larger than the 256 MB Infinity Cache: the practical
    HBM ceiling beside the 8 TB/s peak."""
    import torch

    a = torch.ones(mib * 2 ** 18, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(repeats):
        b.copy_(a)
    end.record()
    torch.cuda.synchronize()
    seconds = start.elapsed_time(end) * 1e-3 / repeats
    del a, b
    return 2 * mib * 2 ** 20 / seconds / 1e9


def log(message):
    """Progress line on stderr (long workloads)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {message}", file=sys.stderr, flush=True)


def _line(metric, value, unit, steps, warmup, ms, config, dtype="f32", **extra):
    line = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}
    line.update(extra)
    print(json.dumps(line), flush=True)


def bench_adjoint(args):
    """Adjoint gradients/s on the 200x200x80 bend (100x100 voxels, 4000 steps,
    checkpoints every 64), with central differences (h = 1e-3) on 16 seeded
    voxels as the accuracy check."""
    import torch

    import __graft_entry__

    pkg = __graft_entry__.package()
    rho = np.random.default_rng(0).uniform(0, 1, (100, 100))
    path = Path(tempfile.gettempdir()) / "wfdtd_bench_bend_200x200x80"
    if not (path / "manifest.json").exists():
        pkg.cases.bend_case(path, rho, shape=(200, 200, 80), steps=4000, pml=10,
                            design=(100, 100), wavelength=1.55e-6)
    every = 64 if args.checkpoint_every is None else (args.checkpoint_every or None)
    # (0: the whole curl history, no recompute)
    problem = pkg.design.TransmissionProblem(path, checkpoint_every=every)
    problem.value_and_grad(rho)  # compiles the graphs
    t0 = time.perf_counter()
    T, grad = problem.value_and_grad(rho)
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t0
    grad = grad.cpu().numpy()
    fd = None
    if args.fd:
        rng = np.random.default_rng(16)
        picks = rng.choice(100 * 100, 16, replace=False)
        rows = []
        for p_ in picks:
            i, j = divmod(int(p_), 100)
            up, dn = rho.copy(), rho.copy()
            up[i, j] += 1e-3
            dn[i, j] -= 1e-3
            rows.append(((problem.forward(up) - problem.forward(dn)) / 2e-3, grad[i, j]))
        f, a = np.array(rows).T
        fd = {"voxels": 16, "h": 1e-3, "rel_l2": float(np.linalg.norm(a - f) / np.linalg.norm(f)),
              "max_rel": float(np.max(np.abs(a - f) / np.abs(f))), "fd": f.tolist(),
              "adjoint": a.tolist(), "precision": "float32 forward runs on the GPU"}
    n = problem.adjoint.n
    memory = (f"checkpoints/{every}" if every else
              "full curl history in HBM, no recompute")
    _line(f"adjoint gradients/s (200x200x80 bend, 100x100 voxels, 4000 steps, {memory})",
          round(1.0 / seconds, 4), "gradients/s", 4000, 1, round(seconds * 1e3 / 4000, 5),
          {"workload": "90-degree Si bend 200x200x80 @ 20 nm, rho~U[0,1) seed 0, "
                       "eps = 2.085 + rho*9.995 through 11 z-cells, TE0 T at 1.55 um",
           "cells": n, "T": T, "forward_s": problem.timings["forward_s"],
           "backward_s": problem.timings["backward_s"],
           "forward_Mcell_updates_per_s": round(n * 4000 / problem.timings["forward_s"] / 1e6,
                                                1)},
          fd_check=fd)


def bench_slab(args):
    """The 64^3 dielectric slab (10-cell CPML, y-polarised Gaussian plane pulse,
    5-wavelength flux DFT planes, 600 steps): whole-job throughput and the
    roofline of the step kernel over the timed run's launches."""
    import torch

    import __graft_entry__

    pkg = __graft_entry__.package()
    path = Path(tempfile.gettempdir()) / "wfdtd_bench_slab64"
    if not (path / "manifest.json").exists():
        pkg.cases.slab_case(path)
    resident = {"auto": "auto", "on": True, "off": False}[args.resident]
    s = pkg.solver.Solver(path, resident=resident, graph_steps=args.graph_steps)
    pkg.detectors.attach(s)
    s.compile()
    steps = s.steps
    s.reset()
    s.run_steps(steps)
    s.reset()
    torch.cuda.synchronize()
    launches = [] if s.resident else None
    t0 = time.perf_counter()
    s.run_steps(steps, synchronize=False, events=launches)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    s.check()
    B, f_cpml = algorithmic_bytes_per_cell(s.shape, 10)
    if s.resident:
        per = np.array([a.elapsed_time(b) * 1e-3 / n for a, b, n in launches])
        weights = np.array([n for _, _, n in launches], float)
        mean_s = float((per * weights).sum() / weights.sum())
    else:
        mean_s = kernel_time_per_step(pkg, s, args.kernel_repeats)[1]
    achieved = s.n * B / mean_s / 1e9
    _line("Mcell-updates/s + % HBM roofline (64^3 slab, 600 steps)",
          round(s.n * steps / elapsed / 1e6, 1), "Mcell-updates/s", steps, 1,
          round(elapsed / steps * 1e3, 5),
          {"workload": "64^3 vacuum + 16-cell eps 12.08 slab @ 50 nm, 10-cell CPML, Gaussian "
                       "plane pulse, flux DFT planes at 5 wavelengths 1.45-1.65 um",
           "cells": s.n, "resident": bool(s.resident),
           "steps_per_launch": s.graph_steps if s.resident else None},
          roofline={"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "kernel_us_per_step_mean": round(mean_s * 1e6, 3),
                    "bytes_per_cell_update": round(B, 2), "f_cpml": round(f_cpml, 4)})


def bench_batch(args):
    """32 uneven bend designs (120x120x64 .. 256x224x96, 3000 steps, 8-wavelength
    flux DFT planes + a modal objective): batched forward + per-design adjoint."""
    import torch

    import __graft_entry__

    pkg = __graft_entry__.package()
    root = Path(tempfile.gettempdir()) / "wfdtd_bench_batch32"
    t0 = time.perf_counter()
    log("writing 32 design bundles")
    cases = pkg.cases.batch_cases(root, count=32, seed=0, steps=3000)
    gen = time.perf_counter() - t0
    log(f"bundles ready ({gen:.0f} s); compiling")
    every = 192 if args.checkpoint_every is None else (args.checkpoint_every or None)
    bt = pkg.design.BatchTransmission([c for c, _ in cases], wavelength=1.55e-6,
                                      checkpoint_every=every)
    bt.log = log
    bt.compile()
    log("batched forward")
    rhos = [r for _, r in cases]
    t0 = time.perf_counter()
    T = bt.forward(rhos)
    fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    bt.backward()
    torch.cuda.synchronize()
    bwd = time.perf_counter() - t0
    updates = bt.cells * bt.steps
    _line("aggregate Mcell-updates/s (32 uneven bend designs, batched forward)",
          round(updates / fwd / 1e6, 1), "Mcell-updates/s", bt.steps, 0,
          round(fwd * 1e3 / bt.steps, 5),
          {"workload": "32 seeded bends 120x120x64..256x224x96 @ 20 nm, uniform or blurred "
                       "densities, 8-wavelength flux DFT + modal planes, 3000 steps",
           "designs": 32, "cells": bt.cells, "checkpoint_every": every,
           "forward_s": round(fwd, 3),
           "adjoint_s": round(bwd, 3), "designs_per_s_forward_adjoint": round(32 / (fwd + bwd), 4),
           "generation_s": round(gen, 1), "T_mean": float(np.mean(T)),
           "x_chunks": sorted({s.x_chunk for s in bt.batch.solvers}),
           "batched_step_us_by_x_chunk": {k: round(v, 1)
                                          for k, v in bt.batch.x_chunk_us.items()},
           "batched_two_step_pass": bt.batch.use_pass, "pass_per_design": bt.batch.pass_each,
           "batched_us_per_step_pass_vs_fused": bt.batch.pass_us})


MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16


def bench_fno(args):
    """Training throughput of the config-5 FNO (100x100 input, width 32, 12
    modes, 4 layers, 9 outputs, batch 16) on synthetic labels: Adam steps of
    forward + backward on the BF16 MFMA GEMMs.  The roofline object counts
    the GEMMs' multiply-adds (2 flops each) per step over the whole step's
    time, so it bounds their rate from below."""
    import torch

    import __graft_entry__

    pkg = __graft_entry__.package()
    sg, fno = pkg.surrogate, pkg.fno
    side, batch = 100, 16
    net = fno.FNO(side, batch=batch, outputs=9, seed=0)
    rng = np.random.default_rng(0)
    n = 16 * batch
    rhos = rng.uniform(0, 1, (n, side, side))
    T = rng.uniform(0.0, 0.1, (n, 9))
    G = rng.normal(0.0, 1e-4, (n, side, side))
    fit = sg._Fitter(net, sg.Normalizer(T, G), 1.0)
    x, tn, gn = fit.tensors(rhos, T, G)
    batches = [torch.arange(k * batch, (k + 1) * batch, device=net.device) for k in range(16)]
    flops = [0]
    plain = fno.FNO.gemm

    def counted(a, a16, sa, b, b16, sb, c, sc, bt, m, nn, k, *rest, **kw):
        flops[0] += 2 * int(bt) * int(m) * int(nn) * int(k)
        return plain(a, a16, sa, b, b16, sb, c, sc, bt, m, nn, k, *rest, **kw)

    fno.FNO.gemm = staticmethod(counted)
    try:
        fit.step(x, tn, gn, batches[0], 1e-4)
    finally:
        fno.FNO.gemm = staticmethod(plain)
    per_step = flops[0]
    for k in range(args.warmup):
        fit.step(x, tn, gn, batches[k % 16], 1e-4)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    steps = max(args.steps, 1) * 10
    a.record()
    for k in range(steps):
        fit.step(x, tn, gn, batches[k % 16], 1e-4)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / steps
    achieved = per_step / (ms * 1e-3) / 1e12
    _line("FNO training samples/s (100x100 density, width 32, 12 modes, BF16 MFMA)",
          round(batch / (ms * 1e-3), 1), "samples/s", steps, args.warmup, round(ms, 4),
          {"workload": "config-5 FNO, batch 16, Adam step = forward + backward + update on "
                       "synthetic densities and labels", "batch": batch,
           "gemm_flops_per_step": per_step},
          dtype="bf16",
          roofline={"bound": "mfma", "achieved": round(achieved, 3),
                    "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(achieved / MFMA_BF16_PEAK_TFLOPS, 6), "traffic": None,
                    "kernel": "wfd_gemm_bf16 (k_gemm) multiply-adds over the whole step"},
          cpu_baseline=None)


def bench_surrogate_label(args):
    """Training labels of the surrogate: designs first .. first + count - 1 of
    the seeded blurred-random family on the 200x200x80 bend at 20 nm (4000
    steps, checkpoints every 64), each by a forward and a checkpointed adjoint
    (T at 8 wavelengths + 1.55 um, dT/drho at 1.55 um), written to --out
    (.npz).  --check-steps first runs that many plain adjoint ascent steps
    from design 0 and reports their verified T (a passive bend keeps T <= 1)."""
    import torch

    import __graft_entry__

    pkg = __graft_entry__.package()
    sg = pkg.surrogate
    path = Path(tempfile.gettempdir()) / "wfdtd_surrogate_bend_200x200x80"
    if not (path / "manifest.json").exists():
        log("writing the surrogate bend bundle")
        sg.surrogate_case(path)
    ev = sg.Evaluator(path, checkpoint_every=64, log=log)
    check = None
    if args.check_steps:
        rho0 = sg.random_designs(1, 100, seed=args.seed)[0]
        out = sg.optimize(ev, rho0, steps=args.check_steps, step=args.opt_step, log=log)
        spec = ev.transmission(out["rho"])
        check = {"steps": args.check_steps, "step": args.opt_step,
                 "T_history": [round(v, 6) for v in out["history"]], "T_final": out["final"],
                 "final_spectrum": [round(float(v), 6) for v in spec],
                 "seconds": round(out["seconds"], 2)}
        log(f"check: T {out['history'][0]:.4f} -> {out['final']:.4f}")
    rhos = sg.random_designs(args.count, 100, seed=args.seed, start=args.first)
    ev.label(rhos[:1])  # compiles the graphs (not timed)
    t0 = time.perf_counter()
    T, G = ev.label(rhos)
    seconds = time.perf_counter() - t0
    torch.cuda.synchronize()
    if args.out:
        Path(args.out).parent.mkdir(parents=True, exist_ok=True)
        np.savez(args.out, seed=args.seed, first=args.first, count=args.count,
                 wavelengths=sg.WAVELENGTHS, spectra=T, grads=G.astype(np.float32),
                 seconds=seconds)
    f = sg.objective_index()
    _line("surrogate training labels/s (200x200x80 bend, forward + checkpointed adjoint)",
          round(args.count / seconds, 4), "labels/s", args.count, 1,
          round(seconds * 1e3 / args.count, 3),
          {"workload": "90-degree Si bend 200x200x80 @ 20 nm, 100x100-voxel blurred-random "
                       "density (sigma 2 voxels), 4000 steps, checkpoints every 64, modal DFT "
                       "planes at 8 wavelengths 1.50-1.60 um + 1.55 um",
           "seed": args.seed, "first": args.first, "count": args.count,
           "T_max_any_wavelength": float(T.max()),
           "T_objective_range": [float(T[:, f].min()), float(T[:, f].max())],
           "grad_rms": float(np.sqrt((G ** 2).mean())), "out": args.out},
          check=check)


def load_labels(directories):
    """Label files written by --workload surrogate-label in one or more
    directories (comma-separated): (designs, spectra (n, 9), dT/drho (n, 100,
    100)), ordered by seed and first index."""
    import __graft_entry__

    sg = __graft_entry__.package().surrogate
    files = sorted(f for d in directories.split(",") if d for f in Path(d).glob("*.npz"))
    parts = []
    for f in files:
        with np.load(f) as d:
            parts.append((int(d["seed"]), int(d["first"]), int(d["count"]), d["spectra"],
                          d["grads"].astype(np.float32), np.asarray(d["wavelengths"])))
    parts.sort(key=lambda p: (p[0], p[1]))
    rhos, spectra, grads = [], [], []
    for seed, first, count, spec, grad, wl in parts:
        if not np.allclose(wl, sg.WAVELENGTHS, rtol=1e-12, atol=0):
            raise ValueError("labels recorded other wavelengths")
        rhos += sg.random_designs(count, grad.shape[-1], seed=seed, start=first)
        spectra.append(spec)
        grads.append(grad)
    return rhos, np.concatenate(spectra), np.concatenate(grads)


def bench_surrogate(args):
    """Surrogate-assisted inverse design on the 200x200x80 bend (100x100
    voxels, 20 nm, 4000 steps): FNO training on --labels, then the plain
    adjoint loop and the surrogate loop (and its ablation without the network)
    from the same seeded design."""
    import torch

    import __graft_entry__

    pkg = __graft_entry__.package()
    sg = pkg.surrogate
    if not args.labels:
        raise SystemExit("--workload surrogate needs --labels (files of --workload surrogate-label)")
    rhos, T, G = load_labels(args.labels)
    held = args.held_out
    train_rho, T_train, G_train = rhos[:-held], T[:-held], G[:-held]
    held_rho, T_held, G_held = rhos[-held:], T[-held:], G[-held:]
    log(f"{len(train_rho)} training labels, {held} held out")
    path = Path(tempfile.gettempdir()) / "wfdtd_surrogate_bend_200x200x80"
    if not (path / "manifest.json").exists():
        log("writing the surrogate bend bundle")
        sg.surrogate_case(path)
    ev = sg.Evaluator(path, checkpoint_every=64, log=log)
    net = pkg.fno.FNO(100, batch=32, outputs=len(sg.WAVELENGTHS), seed=0)
    t0 = time.perf_counter()
    norm, hist = sg.train(net, train_rho, T_train, G_train, epochs=args.epochs, lr=2e-3,
                          log=log)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    net = sg.inference_copy(net)
    f = sg.objective_index()
    pred = np.array([sg.predict(net, norm, r)[0] for r in held_rho])
    t_err = np.abs(pred[:, f] - T_held[:, f])
    # what the network adds to a lagged adjoint gradient: anchored at a held-out
    # design, one 0.05 ascent step away
    gains = []
    for rho_a, g_a in zip(held_rho[:8], G_held[:8]):
        rho_b = sg._ascent(rho_a, g_a, 0.05)
        _, g_b = ev.adjoint(rho_b)
        corrected = g_a + sg.predict(net, norm, rho_b)[1] - sg.predict(net, norm, rho_a)[1]
        scale = np.linalg.norm(g_b)
        gains.append((float(np.linalg.norm(corrected - g_b) / scale),
                      float(np.linalg.norm(g_a - g_b) / scale)))
    rho0 = sg.random_designs(1, 100, seed=7)[0]
    step = args.opt_step
    log("plain adjoint loop")
    plain = sg.optimize(ev, rho0, steps=args.opt_steps, step=step, log=log)
    target = plain["final"] * (1 - 0.005)
    log(f"plain loop: T {plain['history'][0]:.5f} -> {plain['final']:.5f}; surrogate loop")
    sur = sg.optimize(ev, rho0, steps=args.opt_steps, step=step, surrogate=(net, norm),
                      target=target, log=log)
    log("ablation without the network (lagged adjoint gradient)")
    lag = sg.optimize(ev, rho0, steps=args.opt_steps, step=step, surrogate=(net, norm),
                      lagged=True, target=target, log=log)
    torch.cuda.synchronize()
    spec_plain = ev.transmission(plain["rho"])
    spec_sur = ev.transmission(sur["rho"])
    plain_s = plain["seconds"]
    sur_s = sg.time_to(sur, target)
    lag_s = sg.time_to(lag, target)
    per_hour = 0.0 if sur_s is None else 3600.0 / sur_s
    errors = np.array(sur["errors"])

    def summary(out, reached):
        return {"final_T": out["final"], "seconds": round(out["seconds"], 2),
                "seconds_to_target": None if reached is None else round(reached, 2),
                "speedup_vs_plain": None if reached is None else round(plain_s / reached, 3),
                "iterations": len(out["history"]) - 1, "fallbacks": out["fallbacks"],
                "accepted": out["accepted"], "forwards": out["counts"]["forward"],
                "adjoints": out["counts"]["adjoint"],
                "mean_inner_steps": float(np.mean(out["inner_sizes"])) if out["inner_sizes"]
                else 0.0}

    _line("designs/hr at <=0.5% T error (surrogate-assisted loop, FDTD-verified)",
          round(per_hour, 2), "designs/hr", args.opt_steps, 0,
          round((sur_s or sur["seconds"]) * 1e3, 1),
          {"workload": "90-degree Si bend 200x200x80 @ 20 nm, 100x100-voxel density, 4000 "
                       "steps; 4-layer FNO width 32, 12 modes, BF16 MFMA / FP32 accumulate; "
                       "a design counts once its FDTD-verified T is within 0.5% of the plain "
                       "adjoint loop's final T after 200 steps",
           "training_designs": len(train_rho), "held_out": held, "epochs": args.epochs,
           "train_s": round(train_s, 1), "final_loss": hist[-1],
           "heldout_T_abs_err_median": float(np.median(t_err)),
           "heldout_T_range": [float(T_held[:, f].min()), float(T_held[:, f].max())],
           "gradient_error_network_vs_lagged": [round(float(np.mean([g[0] for g in gains])), 4),
                                                round(float(np.mean([g[1] for g in gains])), 4)],
           "step": step, "target_T": target,
           "plain": {"final_T": plain["final"], "seconds": round(plain_s, 2),
                     "designs_per_hr": round(3600.0 / plain_s, 2),
                     "spectrum_max": float(spec_plain.max())},
           "surrogate": summary(sur, sur_s), "lagged_ablation": summary(lag, lag_s),
           "surrogate_spectrum_max": float(spec_sur.max()),
           "proposal_T_rel_err_median": float(np.median(errors)) if len(errors) else None,
           "proposals_within_1pct": float(np.mean(errors <= 0.01)) if len(errors) else None},
          dtype="bf16")


def main():
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10,
                   help="timed runs of the whole 4000-time-step workload")
    p.add_argument("--warmup", type=int, default=2, help="untimed runs before the timed ones")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle sample")
    p.add_argument("--kernel-repeats", type=int, default=200)
    p.add_argument("--graph-steps", type=int, default=1000,
                   help="steps per captured graph / resident launch (and detector stage slots; "
                        "the library default is the reference's 64)")
    p.add_argument("--x-chunk", type=int, default=0,
                   help="planes per thread column of the fused step (0: autotune)")
    p.add_argument("--resident", choices=("auto", "on", "off"), default="auto",
                   help="on-chip-resident stepping (auto: the faster of the two)")
    p.add_argument("--copy-mib", type=int, default=0,
                   help="only run 8 device-to-device copies of this many MiB (the reference "
                        "traffic for calibrating HBM counters) and exit")
    p.add_argument("--workload", choices=("waveguide", "slab", "bend-adjoint", "batch",
                                           "surrogate", "surrogate-label", "fno"),
                   default="waveguide",
                   help="waveguide: the headline Yee-step metric; the others measure the "
                        "adjoint, batched-design and surrogate workloads")
    p.add_argument("--fd", action="store_true", help="bend-adjoint: central-difference check")
    p.add_argument("--checkpoint-every", type=int, default=None,
                   help="bend-adjoint (default 64) and batch (default 192): steps between "
                        "field checkpoints (0: keep the whole design-box curl history "
                        "instead and skip the recompute)")
    p.add_argument("--labels", default="",
                   help="surrogate: directories of label files (comma-separated)")
    p.add_argument("--held-out", type=int, default=64,
                   help="surrogate: last labels kept out of training")
    p.add_argument("--epochs", type=int, default=80, help="surrogate: training epochs")
    p.add_argument("--opt-steps", type=int, default=200, help="surrogate: optimisation steps")
    p.add_argument("--opt-step", type=float, default=0.01,
                   help="surrogate: largest density change of a step")
    p.add_argument("--seed", type=int, default=0, help="surrogate-label: design family seed")
    p.add_argument("--first", type=int, default=0, help="surrogate-label: first design index")
    p.add_argument("--count", type=int, default=64, help="surrogate-label: designs to label")
    p.add_argument("--out", default="", help="surrogate-label: output .npz")
    p.add_argument("--check-steps", type=int, default=0,
                   help="surrogate-label: plain adjoint ascent steps to run first")
    args = p.parse_args()
    if args.workload != "waveguide" and args.gpus == 1:
        {"slab": bench_slab, "bend-adjoint": bench_adjoint, "batch": bench_batch,
         "surrogate": bench_surrogate, "fno": bench_fno,
         "surrogate-label": bench_surrogate_label}[args.workload](args)
        return
    if args.copy_mib:
        copy_probe(args.copy_mib)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this build "
              "targets a single MI355X; multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        p.error("--steps must be >= 1, --warmup >= 0")
    import torch

    pkg, case = build_case()
    resident = {"auto": "auto", "on": True, "off": False}[args.resident]
    s = pkg.solver.Solver(case, x_chunk=args.x_chunk, resident=resident,
                          graph_steps=args.graph_steps)
    pkg.detectors.attach(s)
    s.compile()
    log(f"compiled: {'resident' if s.resident else 'fused pair'}, {s.steps} steps per run")
    for _ in range(args.warmup):
        s.reset()
        s.run()
    torch.cuda.synchronize()
    start_ev = torch.cuda.Event(enable_timing=True)
    end_ev = torch.cuda.Event(enable_timing=True)
    launches = [] if s.resident else None  # HIP events around each resident launch
    t0 = time.perf_counter()
    start_ev.record()
    for _ in range(args.steps):
        s.reset()
        s.run(events=launches)
    end_ev.record()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    s.check()
    gpu_elapsed = start_ev.elapsed_time(end_ev) * 1e-3
    log(f"timed {args.steps} runs in {elapsed:.3f} s")
    cells = s.n
    time_steps = s.steps * args.steps
    value = cells * time_steps / elapsed / 1e6
    finite = bool(np.isfinite(s.E.numpy()).all() and np.isfinite(s.H.numpy()).all())
    T = pkg.detectors.dft_transmission(s)

    B, f_cpml = algorithmic_bytes_per_cell(s.shape, PML)
    if s.resident:
        # the launches of the timed region itself (staging on, replays between them)
        lengths = np.array([n for _, _, n in launches], float)
        seconds = np.array([a.elapsed_time(b) * 1e-3 for a, b, _ in launches])
        per_launch_steps = int(np.median(lengths))
        mean_s = float(seconds.sum() / lengths.sum())
        median_s = float(np.median(seconds / lengths))
        counts = sorted({int(v) for v in lengths})
        kernel = f"k_resident ({len(launches)} launches of {counts} steps)"
        timing = ("HIP events on the launch stream around every resident launch of the timed "
                  "runs (each with its counter kernel)")
    else:
        median_s, mean_s = kernel_time_per_step(pkg, s, args.kernel_repeats)
        per_launch_steps = 1
        kernel = "k_yee_e + k_yee_h (one fused step)"
        timing = "HIP events around isolated fused E + H kernel pairs"
    achieved = cells * B / mean_s / 1e9
    measured = MEASURED_TRAFFIC.get((tuple(s.shape), bool(s.resident), per_launch_steps))
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": None if measured is None else round(measured[0] * per_launch_steps),
                "traffic_source": None if measured is None else
                f"rocprofv3 FETCH_SIZE + WRITE_SIZE per launch of {per_launch_steps} steps, "
                f"{measured[1]}",
                "steps_per_launch": per_launch_steps,
                "algorithmic_bytes_per_launch": round(cells * B * per_launch_steps),
                "kernel": kernel, "kernel_timing": timing,
                "bytes_per_cell_update": round(B, 2), "f_cpml": round(f_cpml, 4),
                "kernel_us_per_step_mean": round(mean_s * 1e6, 3),
                "kernel_us_per_step_median": round(median_s * 1e6, 3),
                "resident": bool(s.resident)}
    ceiling = copy_ceiling()
    roofline["copy_ceiling_GBs"] = round(ceiling, 1)
    roofline["frac_of_copy_ceiling"] = round(achieved / ceiling, 4)
    if s.resident:
        L = s.layout
        roofline["resident_layout"] = {"sub_box": [L.sx, L.sy, L.sz],
                                       "workgroups": [L.gx, L.gy, L.gz]}
    else:
        roofline["x_chunk"] = s.x_chunk
    line = {
        "metric": "Mcell-updates/s + % HBM roofline (fused Yee+CPML step, 160x96x80 waveguide)",
        "value": round(value, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "straight Si waveguide 160x96x80 @ 20 nm, 12-cell CPML, TE0 "
                               "Hann pulse at 1.55 um, 8-wavelength modal+flux DFT planes; "
                               "one bench step = one complete 4000-time-step run",
                   "grid": list(s.shape), "cells": cells, "time_steps_per_run": s.steps,
                   "time_steps": time_steps, "gpu_seconds": round(gpu_elapsed, 6),
                   "finite": finite,
                   "modal_transmission_1p50_1p60um": [round(float(v), 6) for v in T]},
        "roofline": roofline,
        "cpu_baseline": None if args.no_cpu else cpu_baseline(s),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
