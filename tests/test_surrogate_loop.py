"""The surrogate-assisted loop's acceptance logic (warp-fdtd_amd/surrogate.py
optimize) on an analytic stand-in for the solver and the network (no GPU).

T(rho) = sum(w * rho) / n is linear, so its gradient is w / n; the stand-in
network predicts T with a constant bias and the exact gradient.  Anchored at
the last verified design (the gradient and the difference predictors) the
bias cancels and every proposal is accepted; the network's absolute T misses
the 1 % tolerance and every proposal falls back to the adjoint step.  A network whose gradient points downhill is never
accepted when improvement is required.
"""

import numpy as np
import pytest
import torch

N = 16


class _Evaluator:
    def __init__(self, w):
        self.w = w
        self.counts = {"forward": 0, "adjoint": 0, "labels": 0}

    def _spectrum(self, rho):
        spec = np.zeros(8)
        spec[:] = float((self.w * rho).sum()) / rho.size
        return spec

    def transmission(self, rho):
        self.counts["forward"] += 1
        return self._spectrum(rho)

    def adjoint(self, rho):
        self.counts["adjoint"] += 1
        return self._spectrum(rho), self.w / rho.size


@pytest.fixture()
def loop(pkg, monkeypatch):
    sg = pkg.surrogate
    monkeypatch.setattr(torch.cuda, "synchronize", lambda *a, **k: None)
    rng = np.random.default_rng(0)
    w = rng.uniform(0.5, 1.5, size=(N, N))
    rho0 = np.full((N, N), 0.3)

    def run(bias, sign=1.0, **kw):
        ev = _Evaluator(w)
        net = object()

        def predict(net_, norm, rho):
            t = np.full(8, float((w * rho).sum()) / rho.size + bias)
            return t, sign * w / rho.size

        monkeypatch.setattr(sg, "predict", predict)
        out = sg.optimize(ev, rho0, steps=10, step=0.05, surrogate=(net, None), **kw)
        return out, ev

    return run, w, rho0


def test_corrected_prediction_cancels_a_constant_bias(loop):
    """Both anchored predictors are exact on a linear T: the gradient
    (trapezoid) one by construction, the difference one because the bias
    cancels."""
    run, w, rho0 = loop
    for predictor in ("taylor", "difference"):
        out, ev = run(bias=0.05, predictor=predictor, anchored=False)
        assert out["fallbacks"] == 0 and ev.counts["adjoint"] == 0, predictor
        assert ev.counts["forward"] == 11  # the start and every proposal, verified
        assert max(out["errors"]) < 1e-12
        assert all(b > a for a, b in zip(out["history"], out["history"][1:]))
        assert out["final"] == out["history"][-1]


def test_uncorrected_prediction_falls_back(loop):
    run, w, rho0 = loop
    out, ev = run(bias=0.05, predictor="absolute", anchored=False)
    assert out["fallbacks"] == 10 and ev.counts["adjoint"] == 10
    assert min(out["errors"]) > 0.01
    # the adjoint steps still climb
    assert all(b > a for a, b in zip(out["history"], out["history"][1:]))


def test_downhill_proposals_are_replaced(loop):
    run, w, rho0 = loop
    # the T head is exact, the gradient head points downhill: the predicted
    # T is right, so only the improvement rule rejects the proposals
    out, ev = run(bias=0.0, sign=-1.0, predictor="difference", anchored=False)
    assert out["fallbacks"] == 10
    assert all(b > a for a, b in zip(out["history"], out["history"][1:]))
    out, ev = run(bias=0.0, sign=-1.0, predictor="difference", require_improvement=False,
                  anchored=False)
    assert out["fallbacks"] == 0
    assert all(b < a for a, b in zip(out["history"], out["history"][1:]))


def test_gradient_predictor_rejects_a_wrong_gradient(loop):
    """A gradient head pointing downhill predicts a rise where T falls."""
    run, w, rho0 = loop
    out, ev = run(bias=0.0, sign=-1.0, require_improvement=False, anchored=False)
    assert out["fallbacks"] == 10 and min(out["errors"]) > 0.01


def test_anchored_gradient_corrects_the_network(loop):
    """Anchored at an adjoint gradient, the network only supplies the change
    of gradient: a downhill gradient head that is constant contributes none,
    so after the first (adjoint) step every proposal is exact and accepted."""
    run, w, rho0 = loop
    for lagged in (False, True):
        out, ev = run(bias=0.3, sign=-1.0, lagged=lagged)
        assert out["fallbacks"] == 1 and ev.counts["adjoint"] == 1
        assert ev.counts["forward"] == 11
        assert len(out["errors"]) == 9 and max(out["errors"]) < 1e-12
        assert all(b > a for a, b in zip(out["history"], out["history"][1:]))
