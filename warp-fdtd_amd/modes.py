"""Full-vectorial finite-difference eigenmodes of a uniform-grid cross-section.

Fills the role of the FDTDX/tidy3d mode solver that produced the reference
bundles' ``source/_E``, ``_H``, ``_neff`` (baseline/export_case.py:52-191),
written from Maxwell's equations on the same Yee staggering as the time
stepper: for propagation along x with fields ~ exp(i(beta x - omega t)) and
transverse (y, z), eliminating E_x through div(eps E) = 0 gives

  beta^2 Ey = k0^2 ey Ey - Dzb (Dyf Ez - Dzf Ey) + Dyf ex^-1 (Dyb ey Ey + Dzb ez Ez)
  beta^2 Ez = k0^2 ez Ez + Dyb (Dyf Ez - Dzf Ey) + Dzf ex^-1 (Dyb ey Ey + Dzb ez Ez)

with forward/backward differences whose out-of-domain neighbours are zero,
exactly as the stepper pads its fields.  H follows from curl E.  Units are
the stepper's normalised ones (H scaled by the vacuum impedance), and modes
are normalised to 0.5 * sum(Ey Hz - Ez Hy) = 1 over the plane's cells.
"""

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

C0 = 299792458.0


def _diff(n, forward):
    """(f[i+1]-f[i]) or (f[i]-f[i-1]) with zero beyond the ends, unit spacing."""
    main = -np.ones(n) if forward else np.ones(n)
    off = np.ones(n - 1)
    if forward:
        return sp.diags([main, off], [0, 1], shape=(n, n), format="csr")
    return sp.diags([main, -off], [0, -1], shape=(n, n), format="csr")


@dataclass
class Mode:
    """A mode propagating along the solver's first axis with transverse (u, v).
    Components are (prop, u, v): E[0] (imaginary) at (u, v), E[1] at
    (u+1/2, v), E[2] at (u, v+1/2); H[0] (imaginary) at (u+1/2, v+1/2), H[1]
    at (u, v+1/2), H[2] at (u+1/2, v) — the Yee points of a plane normal to
    the propagation axis (H half a cell downstream)."""

    neff: float
    E: np.ndarray  # (3, nu, nv) complex
    H: np.ndarray  # (3, nu, nv) complex
    wavelength: float


def _avg(f, axis, step):
    """Average neighbours along axis: step -1 pairs (i-1, i) -> i, step +1 pairs
    (i, i+1) -> i+1/2; values beyond the plane are zero (as in the stepper)."""
    g = np.moveaxis(np.array(f, copy=True), axis, 0)
    src = np.moveaxis(np.asarray(f), axis, 0)
    if step < 0:
        g[1:] = 0.5 * (src[1:] + src[:-1])
        g[0] = 0.5 * src[0]
    else:
        g[:-1] = 0.5 * (src[:-1] + src[1:])
        g[-1] = 0.5 * src[-1]
    return np.moveaxis(g, 0, axis)


def colocated(mode, axis, spacing, dt=None):
    """E and H of a mode averaged to the E_z point (x, y, z+1/2) exactly as
    detectors.colocate averages stepped fields (detectors.py:26-62), for a
    plane normal to x (axis 0, (u, v) = (y, z)) or to y (axis 1,
    (u, v) = (z, x)).  Averages along the propagation axis become
    cos(beta h/2); the half-step time average of H becomes cos(omega dt/2).
    Returns (3, a, b) arrays in (x, y, z) component order on the plane's own
    index order ((y, z) for axis 0, (x, z) for axis 1)."""
    E, H = mode.E, mode.H
    cx = np.cos(np.pi * mode.neff * spacing / mode.wavelength)
    ct = 1.0 if dt is None else np.cos(np.pi * C0 / mode.wavelength * dt)
    if axis == 0:  # target (u, v+1/2)
        ex = cx * _avg(E[0], 1, +1)
        ey = _avg(_avg(E[1], 0, -1), 1, +1)
        ez = E[2]
        hx = _avg(H[0], 0, -1)
        hy = cx * H[1]
        hz = cx * _avg(_avg(H[2], 0, -1), 1, +1)
        return np.stack([ex, ey, ez]), ct * np.stack([hx, hy, hz])
    if axis == 1:  # (u, v) = (z, x); target (u+1/2, v)
        ez = E[1]
        ex = _avg(_avg(E[2], 1, -1), 0, +1)
        ey = cx * _avg(E[0], 0, +1)
        hz = cx * _avg(_avg(H[1], 1, -1), 0, +1)
        hx = cx * H[2]
        hy = _avg(H[0], 1, -1)
        ec = np.stack([ex, ey, ez])
        hc = ct * np.stack([hx, hy, hz])
        return ec.transpose(0, 2, 1), hc.transpose(0, 2, 1)
    raise ValueError("mode planes normal to x or y only")


def solve_plane(eps_plane, axis, spacing, wavelength, polarization="TE"):
    """Fundamental mode of a plane normal to axis 0 (eps (ny, nz)) or axis 1
    (eps (nx, nz)); TE means the in-plane horizontal E dominates."""
    if axis == 0:
        return solve(eps_plane, spacing, wavelength, polarization=polarization)[0]
    if axis == 1:
        # transverse (u, v) = (z, x): the horizontal in-plane component is E_v
        pol = {"TE": "TM", "TM": "TE", None: None}[polarization]
        return solve(np.asarray(eps_plane).T, spacing, wavelength, polarization=pol)[0]
    raise ValueError("mode planes normal to x or y only")


def solve(eps, spacing, wavelength, count=1, polarization=None, neff_guess=None):
    """Modes of a cross-section with cell permittivity eps (ny, nz).

    Returns ``count`` modes sorted by decreasing effective index; with
    ``polarization`` "TE" (dominant Ey) or "TM" (dominant Ez) only modes of that
    kind are returned.  The solve is deterministic (fixed ARPACK start vector).
    """
    eps = np.asarray(eps, np.float64)
    ny, nz = eps.shape
    N = ny * nz
    h = float(spacing)
    k0 = 2 * np.pi / wavelength
    Iy, Iz = sp.identity(ny, format="csr"), sp.identity(nz, format="csr")
    Dyf = sp.kron(_diff(ny, True), Iz, format="csr") / h
    Dyb = sp.kron(_diff(ny, False), Iz, format="csr") / h
    Dzf = sp.kron(Iy, _diff(nz, True), format="csr") / h
    Dzb = sp.kron(Iy, _diff(nz, False), format="csr") / h
    e = sp.diags(eps.ravel())
    einv = sp.diags(1.0 / eps.ravel())
    A11 = k0 ** 2 * e + Dzb @ Dzf + Dyf @ einv @ Dyb @ e
    A12 = -Dzb @ Dyf + Dyf @ einv @ Dzb @ e
    A21 = -Dyb @ Dzf + Dzf @ einv @ Dyb @ e
    A22 = k0 ** 2 * e + Dyb @ Dyf + Dzf @ einv @ Dzb @ e
    A = sp.bmat([[A11, A12], [A21, A22]], format="csc")
    nmax = float(np.sqrt(eps.max())) if neff_guess is None else float(neff_guess)
    sigma = (k0 * nmax) ** 2
    want = max(count * 4, 6)
    v0 = np.ones(2 * N) / np.sqrt(2 * N)
    vals, vecs = spla.eigs(A, k=min(want, 2 * N - 2), sigma=sigma, v0=v0)
    order = np.argsort(-vals.real)
    modes = []
    for idx in order:
        b2 = vals[idx].real
        if b2 <= 0:
            continue
        beta = np.sqrt(b2)
        v = vecs[:, idx]
        v = v / v[np.argmax(np.abs(v))]  # real up to rounding for lossless guides
        Ey = v[:N].real.reshape(ny, nz)
        Ez = v[N:].real.reshape(ny, nz)
        frac_y = np.sum(Ey ** 2) / max(np.sum(Ey ** 2) + np.sum(Ez ** 2), 1e-300)
        if polarization == "TE" and frac_y < 0.5:
            continue
        if polarization == "TM" and frac_y >= 0.5:
            continue
        ey, ez = Ey.ravel(), Ez.ravel()
        div = Dyb @ (eps.ravel() * ey) + Dzb @ (eps.ravel() * ez)
        Ex_over_i = div / (beta * eps.ravel())  # Ex = i * this
        Hx_times_i = -(Dyf @ ez - Dzf @ ey) / k0  # Hx = (Dyf Ez - Dzf Ey)/(i k0)
        Hy = (-beta * ez + Dzf @ Ex_over_i) / k0
        Hz = (beta * ey - Dyf @ Ex_over_i) / k0
        E = np.stack([1j * Ex_over_i, ey, ez]).reshape(3, ny, nz)
        H = np.stack([1j * Hx_times_i, Hy, Hz]).reshape(3, ny, nz)
        power = 0.5 * np.sum(E[1].real * H[2].real - E[2].real * H[1].real)
        scale = 1 / np.sqrt(abs(power))
        modes.append(Mode(float(beta / k0), E * scale, H * scale * np.sign(power), wavelength))
        if len(modes) == count:
            break
    if len(modes) < count:
        raise ValueError("Mode solver found fewer guided modes than requested")
    return modes
