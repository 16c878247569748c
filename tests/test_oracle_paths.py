"""The CPU oracle's adjoint, packed-CPML and DrivenUniform restatements pinned
by the reference's invariants, closed forms and fixtures."""

import json

import numpy as np
import pytest

from conftest import EXPECTED, rel_l2
from oracle import kernels as OK
from oracle.model import pack_coefficients


def _coeffs(shape, rows_e, rows_h):
    ce = np.zeros((3, max(shape), 3), np.float32)
    ce[:, :, 1:] = 1
    ch = ce.copy()
    for axis, length in enumerate(shape):
        ce[axis, [0, length - 1]] = rows_e
        ch[axis, [0, length - 1]] = rows_h
    return ce, ch


def test_oracle_cpml_step_transpose():
    """test_adjoint.py:11-57 on the oracle (seed 71, tol 2e-6)."""
    rng = np.random.default_rng(71)
    shape = (6, 7, 8)
    n = int(np.prod(shape))
    initial = [rng.normal(size=(c, n)).astype(np.float32) * 0.1 for c in (3, 3, 6, 6)]
    cot = [rng.normal(size=a.shape).astype(np.float32) * 0.1 for a in initial]
    e, h, pe, ph = [a.copy() for a in initial]
    qe, qh, qpe, qph = [a.copy() for a in cot]
    eps = rng.uniform(0.2, 1, n).astype(np.float32)
    ce, ch = _coeffs(shape, [-0.1, 0.8, 1.2], [-0.13, 0.75, 1.1])
    step = np.zeros(1, np.int32)
    dummy = np.zeros((3, n), np.float32)
    OK.update(True, e, h, eps, pe, ce, dummy, dummy, step, dummy, 1, shape, -1, 0.4, 0)
    OK.update(False, h, e, eps, ph, ch, dummy, dummy, step, dummy, 1, shape, -1, 0.4, 0)
    d = np.zeros((6, n), np.float32)
    hist = np.zeros((1, 3, 1), np.float32)
    grad = np.zeros(1, np.float32)
    OK.prepare(qh, eps, qph, ch, d, shape, 0.4, 0)
    OK.gather(qe, d, hist, grad, step, shape, (0, 0, 0, 0, 0, 0), 1)
    OK.prepare(qe, eps, qpe, ce, d, shape, 0.4, 1)
    OK.gather(qh, d, hist, grad, step, shape, (0, 0, 0, 0, 0, 0), 0)
    lhs = sum(np.sum(a.astype(float) * b) for a, b in zip([e, h, pe, ph], cot))
    rhs = sum(np.sum(a.astype(float) * b) for a, b in zip(initial, [qe, qh, qpe, qph]))
    assert abs(lhs - rhs) / max(abs(lhs), abs(rhs)) < 2e-6


def test_oracle_transpose_catches_a_planted_sign():
    """The same identity fails when gather's sign is flipped (planted in the
    oracle, CPU only): the check is not vacuous."""
    rng = np.random.default_rng(5)
    shape = (5, 6, 7)
    n = int(np.prod(shape))
    q = rng.normal(size=(3, n)).astype(np.float32)
    d = rng.normal(size=(6, n)).astype(np.float32)
    hist = np.zeros((1, 3, 1), np.float32)
    grad = np.zeros(1, np.float32)
    step = np.zeros(1, np.int32)
    a, b = q.copy(), q.copy()
    OK.gather(a, d, hist, grad, step, shape, (0, 0, 0, 0, 0, 0), 0)
    OK.gather(b, -d, hist, grad, step, shape, (0, 0, 0, 0, 0, 0), 0)
    assert rel_l2(a - q, -(b - q)) < 1e-6 and rel_l2(a, b) > 1e-3


def test_oracle_colocation_transpose():
    """test_adjoint.py:60-127 on the oracle (seed 17, tol 2e-6)."""
    rng = np.random.default_rng(17)
    shape = (6, 7, 8)
    n = int(np.prod(shape))
    e, h, hp = [rng.normal(size=(3, n)).astype(np.float32) for _ in range(3)]
    sy, sz = 7, 8
    count = sy * sz
    step = np.zeros(1, np.int32)
    indices = np.zeros(1, np.int32)
    weights = np.array([[0.3, 0.7]], np.float32)
    ph = np.zeros((6, count, 2), np.float32)
    OK.record(e, h, hp, step, weights, indices, ph, np.zeros(1, np.float32),
              np.ones(count, np.float32), shape, (2, 0, 0), (1, sy, sz), 0, 1, 1)
    source = rng.normal(size=(6, count, 2)).astype(np.float32)
    qe = np.zeros_like(e)
    qh = np.zeros_like(h)
    OK.inject(qe, qh, source, weights, step, shape, (2, 0, 0), (1, sy, sz), 0)
    rhs = np.sum(qe.astype(float) * e) + np.sum(qh.astype(float) * h)
    qh[:] = 0
    OK.inject(qe, qh, source, weights, step, shape, (2, 0, 0), (1, sy, sz), 1)
    rhs += np.sum(qh.astype(float) * hp)
    lhs = np.sum(source[..., 0].astype(float) * ph[..., 0] - source[..., 1].astype(float)
                 * ph[..., 1])
    assert abs(lhs - rhs) / max(abs(lhs), abs(rhs)) < 2e-6


def test_oracle_packed_equals_full_bitwise():
    """test_packed.py:12-91 on the oracle: asymmetric sparse slabs, 20 steps."""
    shape = (7, 8, 9)
    n = int(np.prod(shape))
    rng = np.random.default_rng(8)
    fields = rng.normal(size=(2, 3, n)).astype("f")
    eps = rng.uniform(0.1, 1, n).astype("f")
    states = []
    for packed in (False, True):
        e, h = fields[0].copy(), fields[1].copy()
        hist, coeffs = [], []
        for field in range(2):
            c = np.zeros((3, max(shape), 3), "f")
            c[..., 1:] = 1
            for axis, length in enumerate(shape):
                active = [0, length - 1] if field == 0 else [1, 2, length - 2]
                if axis == 1:
                    active = active[:1]
                c[axis, active] = [-0.12, 0.8, 0.9]
            pc, size = pack_coefficients(c, shape)
            coeffs.append(pc if packed else c)
            hist.append(np.zeros((6, size if packed else n), "f"))
        states.append((e, h, hist, coeffs))
    dummy = np.zeros((3, 1), "f")
    step = np.zeros(1, np.int32)
    for _ in range(20):
        for packed, (e, h, hist, coeffs) in enumerate(states):
            OK.update(True, e, h, eps, hist[0], coeffs[0], dummy, dummy, step, dummy, 1000,
                      shape, -1, 0.2, 0, bool(packed))
            OK.update(False, h, e, eps, hist[1], coeffs[1], dummy, dummy, step, dummy, 1000,
                      shape, -1, 0.2, 0, bool(packed))
    np.testing.assert_array_equal(states[0][0], states[1][0])
    np.testing.assert_array_equal(states[0][1], states[1][1])


def test_oracle_epsilon_normal_closed_form():
    """test_driven_guide.py:123-128: harmonic normal / arithmetic tangential
    averages at a corner-free interface point."""
    got = [OK.epsilon_normal([2.4, 1.4, 1.6], 0.05, [4.8, 3.2, 3.2], [0.4, 0.4], 12.25, 2.25, c)
           for c in range(3)]
    np.testing.assert_allclose(got, [7.25, 2 / (1 / 12.25 + 1 / 2.25), 7.25], rtol=2e-5)


def _observer_weights(d_shapes, spacing, xyz, electric, shapes):
    from itertools import product

    offsets = np.cumsum([0] + [int(np.prod(s)) for s in shapes])
    rows_i, rows_w = [], []
    for c in range(3):
        stag = np.zeros(3) if electric else np.full(3, 0.5)
        stag[c] = 0.5 if electric else 0
        index = xyz / spacing - stag
        base = np.clip(np.floor(index).astype(int), 0, np.array(shapes[c]) - 2)
        frac = index - base
        idx = np.empty((len(xyz), 8), np.int64)
        wts = np.empty((len(xyz), 8))
        for j, corner in enumerate(product((0, 1), repeat=3)):
            p = base + np.array(corner)
            idx[:, j] = offsets[c] + np.ravel_multi_index(tuple(p.T), shapes[c])
            wts[:, j] = np.prod(np.where(np.array(corner, bool), frac, 1 - frac), axis=1)
        rows_i.append(idx)
        rows_w.append(wts.astype(np.float32))
    return np.concatenate(rows_i), np.concatenate(rows_w)


@pytest.mark.slow
def test_oracle_driven_guide_matches_warp_u050():
    """The oracle's DrivenUniform restatement (driven_uniform.py:90-189) with
    the guide observer's time-centred trilinear samples (guide_observer.py:
    37-131) against the Warp samples of octree_guide_controls/u050_small."""
    meta = json.loads((EXPECTED / "u050_small/result.json").read_text())
    want = np.load(EXPECTED / "u050_small/fields.npz")
    p = meta["problem"]
    h = meta["spacing"]
    steps = meta["steps"]
    L = [p["length"], p["width"], p["height"]]
    n = [int(round(x / h)) for x in L]
    dt = meta["dt"]
    eshapes = [tuple(v if a == c else v + 1 for a, v in enumerate(n)) for c in range(3)]
    hshapes = [tuple(v + 1 if a == c else v for a, v in enumerate(n)) for c in range(3)]
    e = np.zeros(sum(int(np.prod(s)) for s in eshapes), np.float32)
    hf = np.zeros(sum(int(np.prod(s)) for s in hshapes), np.float32)
    t = (np.arange(steps + 64) + 0.5) * dt
    wave = (np.exp(-0.5 * ((t - p["pulse_center"]) / p["pulse_sigma"]) ** 2)
            * np.sin(2 * np.pi * t / p["wavelength"])).astype(np.float32)
    counter = np.zeros(1, np.int32)
    xyz = want["xyz"]
    ei, ew = _observer_weights(None, h, xyz, True, eshapes)
    hi, hw = _observer_weights(None, h, xyz, False, hshapes)
    args = (n, h, dt, L, p["sponge_width"], p["sponge_rate"])
    src = [p["source_x"], p["width"] / 2, p["height"] / 2]
    sig = [p["source_sigma_x"], p["source_sigma_transverse"], p["source_sigma_transverse"]]
    previous = 0
    for k, checkpoint in enumerate((3 * steps // 7, 5 * steps // 7, steps)):
        for _ in range(checkpoint - previous):
            OK.driven_magnetic(e, hf, *args)
            OK.driven_electric(hf, e, n, h, dt, L, [p["core_width"], p["core_height"]],
                               p["core_epsilon"], p["clad_epsilon"], p["sponge_width"],
                               p["sponge_rate"], src, sig, wave, counter, 0)
            counter[0] += 1
        previous = checkpoint
        old = hf.copy()
        centred = hf.copy()
        OK.driven_magnetic(e, centred, *args)
        esamp = np.sum(ew * e[ei], axis=1).reshape(3, -1)
        hsamp = np.sum(hw * (0.5 * (centred[hi] + old[hi])), axis=1).reshape(3, -1)
        got = np.concatenate([esamp, hsamp])
        assert rel_l2(got, want["fields"][k]) < 1e-4


def _plane_fields(seed, shape):
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    return [rng.normal(size=(3, n)).astype(np.float32) for _ in range(3)]


@pytest.mark.parametrize("exact", [0, 1])
def test_oracle_record_dft_single_frequency_equals_record(exact):
    """The multi-frequency FP64 DFT with one frequency equals the reference's
    single-frequency phasor record (detectors.py:65-106) on the same fields
    and weights, up to the float32 accumulation of the latter; a swapped
    real/imaginary weight or a wrong component order fails."""
    shape, lo, sub = (6, 7, 8), (2, 1, 0), (1, 6, 8)
    count = int(np.prod(sub))
    rng = np.random.default_rng(5)
    steps = 12
    w32 = rng.normal(size=(steps, 2)).astype(np.float32)
    ph32 = np.zeros((6, count, 2), np.float32)
    ph64 = np.zeros((1, 6, count, 2), np.float64)
    indices = np.arange(steps, dtype=np.int32)
    for t in range(steps):
        e, h, hp = _plane_fields(100 + t, shape)
        OK.record(e, h, hp, np.array([t], np.int32), w32, indices, ph32,
                  np.zeros(1, np.float32), np.ones(count, np.float32), shape, lo, sub, 0,
                  exact, 1)
        OK.record_dft(e, h, hp, t, w32.astype(np.float64)[:, None, :], ph64, shape, lo, sub,
                      exact)
    assert rel_l2(ph64[0], ph32) < 2e-6
    swapped = ph64[0][..., ::-1]
    assert rel_l2(swapped, ph32) > 0.1
    assert rel_l2(ph64[0][[1, 0, 2, 3, 4, 5]], ph32) > 0.1


def test_oracle_record_dft_is_linear_per_frequency():
    """Each frequency's accumulator depends only on its own weights: three
    frequencies at once equal three single-frequency runs bit for bit, and
    static fields accumulate to (sum of weights) x (the colocated value)."""
    shape, lo, sub = (5, 6, 7), (1, 0, 0), (1, 6, 7)
    count = int(np.prod(sub))
    rng = np.random.default_rng(9)
    steps, nf = 9, 3
    w = rng.normal(size=(steps, nf, 2))
    e, h, hp = _plane_fields(3, shape)
    joint = np.zeros((nf, 6, count, 2))
    single = [np.zeros((1, 6, count, 2)) for _ in range(nf)]
    for t in range(steps):
        OK.record_dft(e, h, hp, t, w, joint, shape, lo, sub, 1)
        for f in range(nf):
            OK.record_dft(e, h, hp, t, np.ascontiguousarray(w[:, f:f + 1]), single[f], shape,
                          lo, sub, 1)
    for f in range(nf):
        np.testing.assert_array_equal(joint[f], single[f][0])
    # static fields: one step with weight (1, 0) gives the colocated values
    unit = np.zeros((1, 6, count, 2))
    OK.record_dft(e, h, hp, 0, np.array([[[1.0, 0.0]]]), unit, shape, lo, sub, 1)
    v = unit[0, :, :, 0]
    total = w.sum(axis=0)  # (nf, 2)
    want = v[None, :, :, None] * total[:, None, None, :]
    assert rel_l2(joint, want) < 1e-13


def test_frequency_estimator_closed_form(pkg):
    """The dispersion check's estimator recovers the frequency and decay of
    synthetic single-frequency traces (any phase and offset) to rounding,
    and refuses a trace that does not oscillate."""
    PERIOD, fit_frequency = pkg.convergence.PERIOD, pkg.convergence.fit_frequency

    dt = PERIOD / 41.7
    t = np.arange(1, 1501) * dt
    for ratio, phase, offset in ((0.95, 0.3, 0.0), (1.0, -1.2, 0.2), (1.04, 2.0, -0.1)):
        s = np.cos(2 * np.pi * ratio * t / PERIOD + phase) + offset
        got, residual, decay = fit_frequency(s, dt)
        assert abs(got / ratio - 1) < 1e-12 and residual < 1e-10 and decay == 0
        damped = np.exp(-0.03 * t / PERIOD) * (s - offset) + offset
        got, residual, decay = fit_frequency(damped, dt, damped=True)
        assert abs(got / ratio - 1) < 1e-10 and abs(decay - 0.03) < 1e-9
    with pytest.raises(ValueError):
        fit_frequency(np.exp(t / PERIOD), dt)
