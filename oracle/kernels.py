"""ctypes binding of the C++ CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may use
this module, and only as the checker (or the timed CPU baseline).  Every entry
point restates a reference Warp kernel; see wfdtd_oracle.cpp for the citations.
Arrays are NumPy, C-contiguous, float32 (``f32``) or float64 (``f64``).
"""

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libwfdtd_oracle.so"
_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_L = ctypes.c_int64


def build():
    """Compile the oracle with its Makefile (g++, OpenMP)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
        _lib.ora_set_threads.argtypes = [_I]
        _lib.ora_set_threads.restype = _I
        for suf, real in (("f32", ctypes.c_float), ("f64", ctypes.c_double)):
            R = real
            sig = {
                "update": [_I, _I, _P, _P, _P, _P, _L, _P, _I, _P, _L, _P, _L, _P, _P,
                           _I, _I, _I, _I, _I, R, _I],
                "record": [_P, _P, _P, _P, _P, _P, _P, _P, _P] + [_I] * 12,
                "record_dft": [_P, _P, _P, _I, _P, _I, _P] + [_I] * 10,
                "record_curl": [_P, _P, _P] + [_I] * 9 + [R],
                "prepare": [_P, _P, _P, _P, _I, _P, _I, _I, _I, R, _I],
                "gather": [_P, _P, _P, _P, _P] + [_I] * 10,
                "gather_f64acc": [_P, _P, _P, _P, _P] + [_I] * 10,
                "inject": [_P, _P, _P, _P, _P] + [_I] * 8 + [_L, _I],
                "write_design": [_P, _P] + [_I] * 9,
                "driven_magnetic": [_P, _P, _P, R, R, _P, R, R],
                "driven_electric": [_P, _P, _P, R, R, _P, _P, R, R, R, R, _P, _P, _P,
                                    _P, _I],
                "epsilon_normal": [_P, R, _P, _P, R, R, _I],
                "record_volume": [_P, _P, _P, _P, _P, _P, _P, _L, _I, _I, _I, _I],
                "material_smooth": [_P, _P, _P, _I, _I],
                "material_subpixel": [_P, _P, _I, _I, R],
            }
            for name, args in sig.items():
                fn = getattr(_lib, f"ora_{name}_{suf}")
                fn.argtypes = args
                fn.restype = R if name == "epsilon_normal" else None
    return _lib


def set_threads(threads):
    """OpenMP threads of the oracle's loops; returns the previous count."""
    return int(lib().ora_set_threads(int(threads)))


def _suffix(a):
    if a.dtype == np.float32:
        return "f32"
    if a.dtype == np.float64:
        return "f64"
    raise TypeError(f"oracle arrays must be float32 or float64, not {a.dtype}")


def _ptr(a):
    if a is None:
        return None
    if not a.flags.c_contiguous:
        raise ValueError("oracle arrays must be C-contiguous")
    return a.ctypes.data


def _fn(name, a):
    return getattr(lib(), f"ora_{name}_{_suffix(a)}")


def update(electric, f, g, inv_eps, psi, coeff, source, waves, step, hprev,
           save_h_from, shape, source_x, courant, periodic, packed=False):
    """kernels.py:77-142.  ``step`` is an int32 array of length >= 1."""
    nx, ny, nz = shape
    _fn("update", f)(
        int(electric), int(packed), _ptr(f), _ptr(g), _ptr(inv_eps), _ptr(psi),
        psi.shape[1], _ptr(coeff), coeff.shape[1], _ptr(source), source.shape[1],
        _ptr(waves), waves.shape[1], _ptr(step), _ptr(hprev), int(save_h_from),
        nx, ny, nz, int(source_x), float(courant), int(periodic))


def record(e, h, prev, step, weights, indices, phasor, flux, area, shape, lo,
           sub, axis, exact, is_phasor):
    """detectors.py:65-106.  weights (steps, 2); phasor (6, count, 2)."""
    _fn("record", e)(
        _ptr(e), _ptr(h), _ptr(prev), _ptr(step), _ptr(weights), _ptr(indices),
        _ptr(phasor), _ptr(flux), _ptr(area), *shape, *lo, *sub, int(axis),
        int(exact), int(is_phasor))


def record_dft(e, h, prev, t, weights, phasor, shape, lo, sub, exact=True):
    """Multi-frequency DFT; weights (steps, nf, 2) float64, phasor (nf, 6, count, 2)."""
    nf = weights.shape[1]
    _fn("record_dft", e)(
        _ptr(e), _ptr(h), _ptr(prev), int(t), _ptr(weights), nf, _ptr(phasor),
        *shape, *lo, *sub, int(exact))


def record_curl(h, hist, step, shape, box, courant):
    """adjoint.py:17-43"""
    _fn("record_curl", h)(_ptr(h), _ptr(hist), _ptr(step), *shape, *box, float(courant))


def prepare(q, eps, psi, coeff, d, shape, courant, electric):
    """adjoint.py:46-88"""
    _fn("prepare", q)(_ptr(q), _ptr(eps), _ptr(psi), _ptr(coeff), coeff.shape[1],
                      _ptr(d), *shape, float(courant), int(electric))


def gather(q, d, hist, grad, step, shape, box, from_h):
    """adjoint.py:107-162; a float64 ``grad`` selects FP64 accumulation."""
    name = "gather_f64acc" if (grad.dtype == np.float64 and q.dtype != np.float64) else "gather"
    _fn(name, q)(_ptr(q), _ptr(d), _ptr(hist), _ptr(grad), _ptr(step), *shape, *box,
                 int(from_h))


def inject(qe, qh, source, weights, step, shape, lo, sub, old_h_only):
    """adjoint.py:181-223.  source (6, count, 2); weights (steps, 2)."""
    count = source.shape[1]
    _fn("inject", qe)(_ptr(qe), _ptr(qh), _ptr(source), _ptr(weights), _ptr(step),
                      *shape, *lo, *sub[1:], count, int(old_h_only))


def write_design(plane, eps, shape, box):
    """adjoint.py:472-490"""
    _fn("write_design", eps)(_ptr(plane), _ptr(eps), *shape, *box)


def driven_magnetic(e, h, n, spacing, dt, lengths, width, rate):
    """driven_uniform.py:90-130"""
    real = e.dtype.type
    n_arr = np.ascontiguousarray(n, np.int32)  # held until the call returns
    len_arr = np.ascontiguousarray(lengths, e.dtype)
    _fn("driven_magnetic", e)(
        _ptr(e), _ptr(h), _ptr(n_arr), float(real(spacing)), float(real(dt)), _ptr(len_arr),
        float(real(width)), float(real(rate)))


def driven_electric(h, e, n, spacing, dt, lengths, core, eps_core, eps_clad, width,
                    rate, src, src_sigma, wave, counter, normal_average):
    """driven_uniform.py:133-189"""
    real = e.dtype.type
    keep = [np.ascontiguousarray(n, np.int32), np.ascontiguousarray(lengths, e.dtype),
            np.ascontiguousarray(core, e.dtype), np.ascontiguousarray(src, e.dtype),
            np.ascontiguousarray(src_sigma, e.dtype)]
    _fn("driven_electric", e)(
        _ptr(h), _ptr(e), _ptr(keep[0]), float(real(spacing)), float(real(dt)),
        _ptr(keep[1]), _ptr(keep[2]), float(real(eps_core)), float(real(eps_clad)),
        float(real(width)), float(real(rate)), _ptr(keep[3]), _ptr(keep[4]), _ptr(wave),
        _ptr(counter), int(normal_average))


def epsilon_normal(p, spacing, lengths, core, eps_core, eps_clad, component,
                   dtype=np.float32):
    """driven_uniform.py:62-87"""
    real = np.dtype(dtype).type
    keep = [np.ascontiguousarray(p, dtype), np.ascontiguousarray(lengths, dtype),
            np.ascontiguousarray(core, dtype)]
    return getattr(lib(), "ora_epsilon_normal_" + _suffix(np.zeros(1, dtype)))(
        _ptr(keep[0]), float(real(spacing)), _ptr(keep[1]), _ptr(keep[2]),
        float(real(eps_core)), float(real(eps_clad)), int(component))


def record_volume(e, h, prev, positions, weights, step, phasor, shape, first):
    """volume.py:274-302.  positions (count, 3); phasor (6, count, 2)."""
    _fn("record_volume", e)(_ptr(e), _ptr(h), _ptr(prev), _ptr(positions),
                            _ptr(weights), _ptr(step), _ptr(phasor),
                            positions.shape[0], *shape, int(first))


def material_smooth(p, weights, out):
    """material.py:12-25"""
    _fn("material_smooth", p)(_ptr(p), _ptr(weights), _ptr(out), *p.shape)


def material_subpixel(rho, out, beta):
    """material.py:35-55"""
    real = rho.dtype.type
    _fn("material_subpixel", rho)(_ptr(rho), _ptr(out), *rho.shape, float(real(beta)))
