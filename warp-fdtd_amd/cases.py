"""Synthetic input bundles in the reference's format (manifest.json + inputs.npz).

Mirrors what baseline/export_case.py:76-190 writes for FDTDX scenes, so that
every synthetic workload round-trips through ``Solver(case)``:

* CPML rows from the Roden-Gedney grading (cubic sigma, linear alpha,
  kappa = 1), with the cell offsets of the reference bundles (E at the cell,
  H half a cell inward; pinned against every bundle slab in the tests);
* plane or modal sources with single-frequency (reference), Gaussian-pulse
  or Hann-pulse temporal profiles;
* reference flux/mode detectors and multi-wavelength DFT planes.

Workloads: ``slab_case`` (64^3 vacuum + dielectric slab, plane Gaussian
pulse, 5-wavelength DFT flux planes), ``waveguide_case`` (160x96x80 straight
Si guide at 20 nm, TE0 modal Hann pulse, 8-wavelength modal + flux DFT
planes) and ``bend_case`` (200x200x80 bend with a 100x100 design region).
"""

import hashlib
import json
import os
from pathlib import Path

import numpy as np

from . import modes

C0 = 299792458.0
EPS0 = 8.8541878128e-12
ETA0 = 376.730313668
EPS_SI = 12.08
EPS_OXIDE = 2.085


# ------------------------------------------------------------------ CPML
def cpml_rows(L, spacing, dt, direction, field, frequency, order=3, reflection=1e-6,
              alpha_fraction=0.01):
    """(a, b, 1/kappa) of an L-cell slab; cell 0 is the first cell of the slab
    in increasing coordinate.  sigma_max = -(m+1) ln R / (2 eta0 L dx),
    alpha_max = alpha_fraction * omega0 * eps0, b = exp(-(sigma+alpha) dt/eps0),
    a = sigma/(sigma+alpha) (b-1)."""
    i = np.arange(L, dtype=np.float64)
    if direction == "-":
        depth = (L - i - (1.0 if field == "E" else 1.5)) / L
    else:
        depth = (i - (0.5 if field == "E" else 0.0)) / L
    depth = np.clip(depth, 0, None)
    sigma_max = -(order + 1) * np.log(reflection) / (2 * ETA0 * L * spacing)
    alpha_max = alpha_fraction * 2 * np.pi * frequency * EPS0
    sigma = sigma_max * depth ** order
    alpha = alpha_max * (1 - depth)
    b = np.exp(-(sigma + alpha) * dt / EPS0)
    a = sigma / (sigma + alpha) * (b - 1)
    return a.astype(np.float32), b.astype(np.float32), np.ones(L, np.float32), sigma_max, alpha_max


def add_cpml(objects, payload, shape, L, spacing, dt, frequency):
    k = 0
    for axis in range(3):
        for direction in "-+":
            lo, hi = (0, L) if direction == "-" else (shape[axis] - L, shape[axis])
            grid = [[0, n] for n in shape]
            grid[axis] = [lo, hi]
            name = f"Object_{k}"
            dims = [1, 1, 1]
            dims[axis] = L
            for field in "EH":
                a, b, ik, smax, amax = cpml_rows(L, spacing, dt, direction, field, frequency)
                payload[f"{name}/pml_a_{field}"] = a.reshape(dims)
                payload[f"{name}/pml_b_{field}"] = b.reshape(dims)
                payload[f"{name}/inv_kappa_{field}"] = ik.reshape(dims)
            objects.append({"name": name, "class": "PerfectlyMatchedLayer", "grid_slice": grid,
                            "direction": direction, "sigma_end": float(smax),
                            "alpha_start": float(amax), "sigma_order": 3.0,
                            "kappa_start": 1.0, "kappa_end": 1.0})
            k += 1


# --------------------------------------------------------------- writing
def write_bundle(path, meta, payload):
    """inputs.npz + manifest.json with the sha256 the loader checks."""
    path = Path(path)
    path.mkdir(parents=True, exist_ok=True)
    raw = path / "inputs.npz"
    np.savez(raw, **payload)
    meta = dict(meta)
    meta["inputs_sha256"] = hashlib.sha256(raw.read_bytes()).hexdigest()
    meta["arrays"] = {k: {"shape": list(v.shape), "dtype": str(v.dtype)}
                      for k, v in payload.items()}
    (path / "manifest.json").write_text(json.dumps(meta, indent=1))
    return path


def _base(shape, spacing, courant_factor, steps, inv_eps, **settings):
    dt = courant_factor / np.sqrt(3) * spacing / C0
    payload = {"inv_permittivity": np.ascontiguousarray(inv_eps[None], np.float32),
               "inv_permeability": np.array(1.0)}
    for a, axis in enumerate("xyz"):
        payload[f"{axis}_edges_m"] = (np.arange(shape[a] + 1) * spacing).astype(np.float32)
    meta = {"settings": {"spacing_nm": spacing * 1e9, "precision": "float32",
                         "courant_factor": courant_factor, **settings},
            "dt_s": float(dt), "time_steps": int(steps), "objects": [],
            "convention": "normalised fields (H scaled by the vacuum impedance), Yee"}
    return meta, payload, dt


def _source(meta, payload, shape, x, E, H, neff, spacing, dt, profile):
    """Plane source at x with real E/H profiles (3, ny, nz); H lags E by the
    half-cell travel time neff*dx/(2c) (the reference's time offsets)."""
    ny, nz = shape[1:]
    name = "source"
    payload[f"{name}/_E"] = np.ascontiguousarray(E.reshape(3, 1, ny, nz), np.float32)
    payload[f"{name}/_H"] = np.ascontiguousarray(H.reshape(3, 1, ny, nz), np.float32)
    payload[f"{name}/_neff"] = np.array(neff, np.complex64)
    delay = neff * spacing / 2 / C0 / dt
    off_e = np.zeros((3, 1, ny, nz), np.float32)
    off_h = np.zeros((3, 1, ny, nz), np.float32)
    off_e[0] = -delay
    off_h[1:] = -delay
    payload[f"{name}/_time_offset_E"] = off_e
    payload[f"{name}/_time_offset_H"] = off_h
    payload[f"{name}/_time_step_to_on_idx"] = np.arange(meta["time_steps"], dtype=np.int32)
    obj = {"name": name, "class": "ModePlaneSource",
           "grid_slice": [[x, x + 1], [0, ny], [0, nz]], "direction": "+",
           "propagation_axis": 0, "static_amplitude_factor": 1.0, "phase_shift": 0.0,
           "profile_phase_shift": 0.0}
    obj.update(profile)
    meta["objects"].insert(0, obj)


def gaussian_pulse(wavelength, sigma_fraction):
    """Temporal profile: exp(-(t-t0)^2/(2 s^2)) cos(2 pi f (t-t0)), spectral
    sigma = sigma_fraction * f, delay t0 = 4 s."""
    f = C0 / wavelength
    s = 1 / (2 * np.pi * sigma_fraction * f)
    return {"temporal_profile": "GaussianPulseProfile", "frequency_hz": f, "sigma_s": s,
            "delay_s": 4 * s}


def hann_pulse(wavelength, duration):
    """Temporal profile: sin^2(pi t/T) cos(2 pi f t) for 0 <= t <= T."""
    return {"temporal_profile": "HannPulseProfile", "frequency_hz": C0 / wavelength,
            "duration_s": float(duration)}


def _dft_plane(meta, payload, name, grid, axis, wavelengths, mode_list=None, spacing=None,
               dt=None):
    obj = {"name": name, "class": "DFTPlaneDetector", "grid_slice": grid,
           "propagation_axis": axis, "direction": "+",
           "wavelengths_m": [float(w) for w in wavelengths], "exact_interpolation": True}
    meta["objects"].append(obj)
    if mode_list is not None:
        pairs = [modes.colocated(m, axis, spacing, dt) for m in mode_list]
        ec = np.stack([p[0] for p in pairs])
        hc = np.stack([p[1] for p in pairs])
        shape = [hi - lo for lo, hi in grid]
        payload[f"{name}/_mode_E"] = ec.reshape(len(mode_list), 3, *shape).astype(np.complex64)
        payload[f"{name}/_mode_H"] = hc.reshape(len(mode_list), 3, *shape).astype(np.complex64)
        payload[f"{name}/_mode_neff"] = np.array([m.neff for m in mode_list], np.float64)


# ------------------------------------------------------------- workloads
def slab_case(path, n=64, spacing=50e-9, steps=600, pml=10, thickness=16, eps_slab=12.08,
              courant_factor=0.99, wavelengths=(1.45e-6, 1.50e-6, 1.55e-6, 1.60e-6, 1.65e-6),
              sigma_fraction=0.1, vacuum=False):
    """64^3 vacuum with a dielectric slab normal to x, y-polarised Gaussian
    plane-wave pulse, DFT flux planes before and after the slab."""
    shape = (n, n, n)
    eps = np.ones(shape)
    x0 = (n - thickness) // 2
    if not vacuum:
        eps[x0:x0 + thickness] = eps_slab
    meta, payload, dt = _base(shape, spacing, courant_factor, steps, 1 / eps, scene="slab",
                              pml_cells=pml, slab_cells=thickness, eps_slab=eps_slab)
    f0 = C0 / 1.55e-6
    add_cpml(meta["objects"], payload, shape, pml, spacing, dt, f0)
    E = np.zeros((3, n, n))
    H = np.zeros((3, n, n))
    E[1] = 1.0  # y-polarised, vacuum plane wave: Hz = Ey in normalised units
    H[2] = 1.0
    sx = pml + 2
    _source(meta, payload, shape, sx, E, H, 1.0, spacing, dt,
            gaussian_pulse(1.55e-6, sigma_fraction))
    inner = [pml + 1, n - pml - 1]
    _dft_plane(meta, payload, "flux_before", [[sx + 3, sx + 4], inner, inner], 0, wavelengths)
    xa = x0 + thickness + 4
    _dft_plane(meta, payload, "flux_after", [[xa, xa + 1], inner, inner], 0, wavelengths)
    return write_bundle(path, meta, payload)


def waveguide_cross_section(ny, nz, spacing, width=500e-9, height=220e-9, core=EPS_SI,
                            clad=EPS_OXIDE):
    """Cell permittivity of a centred rectangular core (integer-cell edges)."""
    eps = np.full((ny, nz), clad)
    wy = int(round(width / spacing))
    wz = int(round(height / spacing))
    y0 = (ny - wy) // 2
    z0 = (nz - wz) // 2
    eps[y0:y0 + wy, z0:z0 + wz] = core
    return eps


def waveguide_case(path, shape=(160, 96, 80), spacing=20e-9, steps=4000, pml=12,
                   courant_factor=0.99, wavelengths=None, pulse_duration=60e-15,
                   center_wavelength=1.55e-6):
    """Straight 500x220 nm Si guide along x in oxide, TE0 modal source with a
    Hann-windowed pulse, modal + flux DFT planes at 8 wavelengths."""
    if wavelengths is None:
        wavelengths = np.linspace(1.50e-6, 1.60e-6, 8)
    nx, ny, nz = shape
    section = waveguide_cross_section(ny, nz, spacing)
    eps = np.broadcast_to(section, shape).copy()
    meta, payload, dt = _base(shape, spacing, courant_factor, steps, 1 / eps,
                              scene="straight_waveguide", pml_cells=pml)
    add_cpml(meta["objects"], payload, shape, pml, spacing, dt, C0 / center_wavelength)
    mode = modes.solve(section, spacing, center_wavelength, polarization="TE")[0]
    sx = pml + 3
    _source(meta, payload, shape, sx, mode.E.real, mode.H.real, mode.neff, spacing, dt,
            hann_pulse(center_wavelength, pulse_duration))
    per_wl = [modes.solve(section, spacing, w, polarization="TE")[0] for w in wavelengths]
    full = [[0, ny], [0, nz]]
    xi = sx + 8
    xo = nx - pml - 8
    _dft_plane(meta, payload, "mode_in", [[xi, xi + 1], *full], 0, wavelengths, per_wl,
               spacing, dt)
    _dft_plane(meta, payload, "mode_out", [[xo, xo + 1], *full], 0, wavelengths, per_wl,
               spacing, dt)
    return write_bundle(path, meta, payload)


def bend_case(path, rho, shape=(200, 200, 80), spacing=20e-9, steps=4000, pml=10,
              courant_factor=0.99, wavelength=1.55e-6, design=(100, 100),
              wavelengths=None, pulse_duration=60e-15, flux_wavelengths=None):
    """90-degree bend: a 500x220 nm guide enters along +x, leaves along +y;
    the 2x2 um design region between them holds eps = 2.085 + rho*(12.08-2.085)
    through the 220 nm layer.  rho has shape ``design``.  Modal DFT planes
    mode_in / mode_out record ``wavelengths`` (default: the source's);
    ``flux_wavelengths`` adds flux-only DFT planes flux_in / flux_out on the
    same planes (no mode solve)."""
    nx, ny, nz = shape
    rho = np.asarray(rho, np.float64)
    if rho.shape != tuple(design):
        raise ValueError(f"rho must have shape {design}")
    section = waveguide_cross_section(ny, nz, spacing)
    wy = int(round(500e-9 / spacing))
    wz = int(round(220e-9 / spacing))
    z0 = (nz - wz) // 2
    sx_, sy_ = design
    bx0 = (nx - sx_) // 2
    by0 = (ny - sy_) // 2
    eps = np.full(shape, EPS_OXIDE)
    yc = by0 + (sy_ - wy) // 2
    eps[:bx0, yc:yc + wy, z0:z0 + wz] = EPS_SI  # input arm along x
    xc = bx0 + (sx_ - wy) // 2
    eps[xc:xc + wy, by0 + sy_:, z0:z0 + wz] = EPS_SI  # output arm along y
    eps[bx0:bx0 + sx_, by0:by0 + sy_, z0:z0 + wz] = (EPS_OXIDE + rho * (EPS_SI - EPS_OXIDE))[
        :, :, None]
    meta, payload, dt = _base(shape, spacing, courant_factor, steps, 1 / eps, scene="bend",
                              pml_cells=pml, design_box=[bx0, by0, z0, sx_, sy_, wz])
    add_cpml(meta["objects"], payload, shape, pml, spacing, dt, C0 / wavelength)
    src_section = eps[pml + 3]
    mode_in = modes.solve(src_section, spacing, wavelength, polarization="TE")[0]
    sx = pml + 3
    _source(meta, payload, shape, sx, mode_in.E.real, mode_in.H.real, mode_in.neff, spacing,
            dt, hann_pulse(wavelength, pulse_duration))
    if wavelengths is None:
        wavelengths = [wavelength]
    ins = [modes.solve_plane(eps[sx + 8], 0, spacing, w) for w in wavelengths]
    yo = ny - pml - 8
    outs = [modes.solve_plane(eps[:, yo, :], 1, spacing, w) for w in wavelengths]
    _dft_plane(meta, payload, "mode_in", [[sx + 8, sx + 9], [0, ny], [0, nz]], 0, wavelengths,
               ins, spacing, dt)
    _dft_plane(meta, payload, "mode_out", [[0, nx], [yo, yo + 1], [0, nz]], 1, wavelengths,
               outs, spacing, dt)
    if flux_wavelengths is not None:
        _dft_plane(meta, payload, "flux_in", [[sx + 8, sx + 9], [0, ny], [0, nz]], 0,
                   flux_wavelengths)
        _dft_plane(meta, payload, "flux_out", [[0, nx], [yo, yo + 1], [0, nz]], 1,
                   flux_wavelengths)
    return write_bundle(path, meta, payload)


def design_density(rng, shape, blurred, sigma=2.0):
    """A seeded density in [0, 1]: uniform, or uniform blurred by a Gaussian of
    ``sigma`` voxels (reflecting edges) and stretched back to [0, 1]."""
    rho = rng.uniform(0, 1, shape)
    if blurred:
        from scipy.ndimage import gaussian_filter

        rho = gaussian_filter(rho, sigma, mode="reflect")
        lo, hi = rho.min(), rho.max()
        rho = (rho - lo) / max(hi - lo, 1e-12)
    return rho


def batch_shapes(count=32, seed=0, lo=(120, 120, 64), hi=(256, 224, 96), pml=10):
    """Seeded uneven grid shapes in [lo, hi] (per axis, multiples of 8) with
    each design's in-plane voxel count: 100, or less where the grid leaves
    under 20 cells between the design region and the CPML."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        shape = tuple(int(8 * rng.integers(a // 8, b // 8 + 1)) for a, b in zip(lo, hi))
        side = int(min(100, min(shape[:2]) - 2 * pml - 40))
        blurred = bool(rng.integers(0, 2))
        out.append((shape, side, blurred, int(rng.integers(0, 2 ** 31))))
    return out


def _batch_member(args):
    """One bundle of the batch, in a worker process: its BLAS runs on one
    thread (the pool already occupies the cores; threaded ARPACK vectors in
    every worker oversubscribe them several times over)."""
    from threadpoolctl import threadpool_limits

    path, shape, side, blurred, seed, steps, pml, wavelength, flux_wavelengths = args
    rho = design_density(np.random.default_rng(seed), (side, side), blurred)
    with threadpool_limits(limits=1):
        bend_case(path, rho, shape=shape, steps=steps, pml=pml, design=(side, side),
                  wavelength=wavelength, flux_wavelengths=flux_wavelengths)
    return str(path), rho


def batch_cases(root, count=32, seed=0, steps=3000, pml=10, wavelength=1.55e-6,
                flux_wavelengths=None, workers=None):
    """The batched-design workload: ``count`` bends of seeded uneven shapes
    (120x120x64 .. 256x224x96), densities uniform or Gaussian-blurred
    (sigma 2 voxels), a modal objective at ``wavelength`` and 8-wavelength
    flux DFT planes (1.50-1.60 um).  Bundles are written in parallel
    processes and reused when present.  Returns [(path, rho)]."""
    from concurrent.futures import ProcessPoolExecutor
    from pathlib import Path

    if flux_wavelengths is None:
        flux_wavelengths = list(np.linspace(1.50e-6, 1.60e-6, 8))
    root = Path(root)
    jobs = []
    for k, (shape, side, blurred, s) in enumerate(batch_shapes(count, seed, pml=pml)):
        name = f"design{k:02d}_{shape[0]}x{shape[1]}x{shape[2]}_{side}_{int(blurred)}"
        jobs.append((root / name, shape, side, blurred, s, steps, pml, wavelength,
                     flux_wavelengths))
    todo = [j for j in jobs if not (j[0] / "manifest.json").exists()]
    if todo:
        workers = workers or min(16, len(todo), os.cpu_count() or 1)
        with ProcessPoolExecutor(max_workers=workers) as pool:
            list(pool.map(_batch_member, todo))
    out = []
    for j in jobs:
        path, shape, side, blurred, s = j[:5]
        rho = design_density(np.random.default_rng(s), (side, side), blurred)
        out.append((path, rho))
    return out
