"""The surrogate-assisted loop's acceptance logic (warp-fdtd_amd/surrogate.py
optimize) on analytic stand-ins for the solver and the network (no GPU).

T(rho) = mean(w * rho) - (b / 2) mean((rho - c)^2) has the gradient
(w - b (rho - c)) / N.  With b = 0 it is linear: the anchored model gradient
is exact whatever constant gradient the network predicts, so every proposal
is exact and accepted.  With b > 0 the trapezoid rule on the exact gradient
is still exact (the gradient is linear along a straight step), so a network
that predicts the exact gradient has every proposal accepted, while the
lagged gradient (the method without the network) mispredicts T and falls
back to the adjoint; the network's loop climbs higher for the same number of
iterations.
"""

import numpy as np
import pytest
import torch

N = 16


class _Problem:
    def __init__(self, w, b, c):
        self.w, self.b, self.c = w, b, c
        self.counts = {"forward": 0, "adjoint": 0, "labels": 0}

    def value(self, rho):
        return float((self.w * rho).mean() - 0.5 * self.b * ((rho - self.c) ** 2).mean())

    def gradient(self, rho):
        return (self.w - self.b * (rho - self.c)) / rho.size

    def _spectrum(self, rho):
        return np.full(9, self.value(rho))

    def transmission(self, rho):
        self.counts["forward"] += 1
        return self._spectrum(rho)

    def adjoint(self, rho):
        self.counts["adjoint"] += 1
        return self._spectrum(rho), self.gradient(rho)


@pytest.fixture()
def loop(pkg, monkeypatch):
    sg = pkg.surrogate
    monkeypatch.setattr(torch.cuda, "synchronize", lambda *a, **k: None)
    rng = np.random.default_rng(0)
    w = rng.uniform(0.5, 1.5, size=(N, N))
    rho0 = np.full((N, N), 0.3)

    def run(b=0.0, network="exact", steps=10, **kw):
        ev = _Problem(w, b, 0.5)

        def predict(net_, norm, rho):
            if network == "exact":
                g = ev.gradient(rho)
            else:  # a constant gradient pointing downhill
                g = -w / rho.size
            return np.full(9, ev.value(rho) + 0.3), g

        monkeypatch.setattr(sg, "predict", predict)
        out = sg.optimize(ev, rho0, steps=steps, step=0.02, surrogate=(object(), None), **kw)
        return out, ev

    return run, w, rho0, sg


def test_linear_objective_every_proposal_exact(loop):
    """On a linear T the anchored gradient is exact with any constant network
    gradient (and so is the lagged one): no fallbacks, one adjoint (the first
    anchor), one forward per iteration, m doubling to its cap."""
    run, w, rho0, sg = loop
    for network, lagged in (("constant", False), ("exact", False), ("exact", True)):
        out, ev = run(b=0.0, network=network, lagged=lagged, max_inner=8)
        assert out["fallbacks"] == 0 and ev.counts["adjoint"] == 1, (network, lagged)
        assert ev.counts["forward"] == 10
        assert max(out["errors"]) < 1e-12
        assert out["inner_sizes"] == [2, 4, 8, 8, 8, 8, 8, 8, 8, 8]
        assert all(b > a for a, b in zip(out["history"], out["history"][1:]))
        assert out["final"] == out["history"][-1] == ev.value(out["rho"])


def test_network_gradient_beats_the_lagged_gradient(loop):
    """On the curved T the exact-gradient network is accepted every time;
    the lagged gradient mispredicts, falls back, and ends lower."""
    run, w, rho0, sg = loop
    net_out, net_ev = run(b=1.0, network="exact", steps=12, max_inner=8)
    lag_out, lag_ev = run(b=1.0, network="exact", steps=12, max_inner=8, lagged=True)
    assert net_out["fallbacks"] == 0 and max(net_out["errors"]) < 1e-9
    assert lag_out["fallbacks"] > 0 and max(lag_out["errors"]) > 0.01
    assert lag_ev.counts["adjoint"] == 1 + lag_out["fallbacks"]
    assert net_out["final"] > lag_out["final"]


def test_target_stops_the_loop(loop):
    run, w, rho0, sg = loop
    full, _ = run(b=0.0, steps=10)
    level = full["history"][3]
    out, ev = run(b=0.0, steps=10, target=level)
    assert out["final"] >= level and len(out["history"]) == 4
    assert sg.time_to(out, level) == out["times"][-1]
    assert sg.time_to(out, out["final"] + 1.0) is None


def test_plain_loop_is_one_adjoint_per_step(loop):
    """The plain loop: one adjoint per step and one forward of the final design."""
    run, w, rho0, sg = loop
    problem = _Problem(w, 0.0, 0.5)
    out = sg.optimize(problem, rho0, steps=7, step=0.02)
    assert problem.counts == {"forward": 1, "adjoint": 7, "labels": 0}
    assert len(out["history"]) == 8 and out["final"] == problem.value(out["rho"])
    assert all(b > a for a, b in zip(out["history"], out["history"][1:]))
