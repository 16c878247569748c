"""Numerical-frequency fit of a recorded field trace (host-side analysis of
warp_fdtd/convergence.py:11-58, used to check Yee dispersion)."""

import numpy as np
from scipy.optimize import least_squares, minimize_scalar

LENGTH = 1550e-9
PERIOD = LENGTH / 299792458.0


def fit_frequency(signal, dt, damped=False):
    """Best-fit frequency ratio (to 1/PERIOD) of a sinusoid plus offset;
    returns (ratio, relative residual, decay)."""
    t = np.arange(1, len(signal) + 1, dtype=float) * dt / PERIOD
    signal = np.asarray(signal, float)

    def loss(ratio):
        phase = 2 * np.pi * ratio * t
        basis = np.column_stack((np.cos(phase), np.sin(phase), np.ones_like(t)))
        coeff = np.linalg.lstsq(basis, signal, rcond=None)[0]
        return np.mean((basis @ coeff - signal) ** 2)

    search = np.linspace(0.9, 1.1, 101)
    i = int(np.argmin([loss(r) for r in search]))
    if not 0 < i < len(search) - 1:
        raise ValueError("frequency outside the 0.9-1.1 search window")
    fit = minimize_scalar(loss, bounds=(search[i - 1], search[i + 1]), method="bounded",
                          options={"xatol": 1e-13})
    if not fit.success:
        raise ValueError("frequency fit failed")
    if damped:
        def residual(v):
            phase = 2 * np.pi * v[0] * t
            envelope = np.exp(-v[1] * t)
            basis = np.column_stack((envelope * np.cos(phase), envelope * np.sin(phase),
                                     np.ones_like(t)))
            return basis @ np.linalg.lstsq(basis, signal, rcond=None)[0] - signal

        joint = least_squares(residual, [fit.x, 0.0], bounds=([0.9, -0.1], [1.1, 1.0]),
                              xtol=1e-13, ftol=1e-13, gtol=1e-13)
        return (float(joint.x[0]), float(np.sqrt(np.mean(joint.fun ** 2) / np.mean(signal ** 2))),
                float(joint.x[1]))
    return float(fit.x), float(np.sqrt(fit.fun / np.mean(signal ** 2))), 0.0
