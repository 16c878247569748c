"""Frequency of a recorded field trace, for the Yee dispersion check of the
periodic propagator (the role of warp_fdtd/convergence.py; this estimator is
a different method).

A sampled sinusoid s_k = A r^k cos(theta k + phi) + c obeys the three-term
linear recurrence (Prony's method / linear prediction)

    s_{k+1} = a1 s_k + a2 s_{k-1} + b,   a1 = 2 r cos(theta),  a2 = -r^2,
                                          b = c (1 - a1 - a2),

exactly, for any amplitude, phase and offset.  Least squares over every
consecutive triple gives (a1, a2, b); theta = acos(a1 / (2 r)) is the phase
advance per sample and -ln(r) the decay per sample.  Undamped traces fix
r = 1 (a2 = -1) and fit a1 and b alone.  For a noise-free single-frequency
trace both are exact up to rounding, with no search window or starting
guess.
"""

import numpy as np

LENGTH = 1550e-9
PERIOD = LENGTH / 299792458.0


def _recurrence(signal, damped):
    s = np.asarray(signal, np.float64)
    if s.ndim != 1 or len(s) < 5:
        raise ValueError("need a 1-D trace of at least 5 samples")
    prev, cur, nxt = s[:-2], s[1:-1], s[2:]
    if damped:
        basis = np.column_stack((cur, prev, np.ones_like(cur)))
        (a1, a2, b), *_ = np.linalg.lstsq(basis, nxt, rcond=None)
    else:
        basis = np.column_stack((cur, np.ones_like(cur)))
        (a1, b), *_ = np.linalg.lstsq(basis, nxt + prev, rcond=None)
        a2 = -1.0
    predicted = a1 * cur + a2 * prev + b
    scale = np.sqrt(np.mean(s * s))
    if not np.isfinite(scale) or scale == 0:
        raise ValueError("trace is zero or not finite")
    residual = float(np.sqrt(np.mean((nxt - predicted) ** 2)) / scale)
    return float(a1), float(a2), residual


def fit_frequency(signal, dt, damped=False):
    """Frequency of the trace (samples dt apart, the first at t = dt) as a
    ratio to 1/PERIOD, the relative RMS residual of the recurrence fit and
    the decay rate per PERIOD (0 unless damped); ValueError when the trace
    is not an oscillation (|cos theta| >= 1)."""
    a1, a2, residual = _recurrence(signal, damped)
    if a2 >= 0:
        raise ValueError("trace does not oscillate (growing or non-oscillatory fit)")
    r = np.sqrt(-a2)
    cos_theta = a1 / (2 * r)
    if not -1 < cos_theta < 1:
        raise ValueError("trace does not oscillate (|cos theta| >= 1)")
    theta = float(np.arccos(cos_theta))
    samples_per_period = PERIOD / dt
    ratio = theta / (2 * np.pi) * samples_per_period
    decay = float(-np.log(r) * samples_per_period) if damped else 0.0
    return ratio, residual, decay
