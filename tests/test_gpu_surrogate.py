"""The FNO surrogate on a small bend: labels, physics of the modal objective,
what the trained network adds to the lagged adjoint gradient, and the
FDTD-verified loop.

A 24x24-voxel bend (64x64x40 cells at 40 nm, 6-cell CPML, 0.44 um of oxide
between the core and the absorber above and below, 700 steps) keeps
labelling, training and the loops within a minute; the code is the 100x100
surrogate workload's.
"""

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIDE = 24


def _say(text):
    print(f"[surrogate] {text}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def setup(tmp_path_factory, pkg):
    sg = pkg.surrogate
    case = sg.surrogate_case(tmp_path_factory.mktemp("surrogate") / "bend", side=SIDE,
                             shape=(64, 64, 40), spacing=40e-9, steps=700, pml=6,
                             pulse_duration=20e-15)
    ev = sg.Evaluator(case, checkpoint_every=64)
    train_rho = sg.random_designs(384, SIDE, seed=1)
    test_rho = sg.random_designs(16, SIDE, seed=2)
    _say("labelling 400 designs")
    T, G = ev.label(train_rho)
    Tt, Gt = ev.label(test_rho)
    net = pkg.fno.FNO(SIDE, batch=16, outputs=len(sg.WAVELENGTHS), seed=0)
    _say("training")
    norm, history = sg.train(net, train_rho, T, G, epochs=60, lr=3e-3)
    return sg, ev, net, norm, history, T, test_rho, Tt, Gt


def test_labels_are_consistent(pkg, gpu, setup):
    """A label equals the single evaluation of the same design; the plain
    forward (verification) gives the label's spectrum."""
    sg, ev, net, norm, history, T, test_rho, Tt, Gt = setup
    spec, g = ev.adjoint(test_rho[0])
    np.testing.assert_array_equal(spec, Tt[0])
    np.testing.assert_array_equal(g, Gt[0])
    fwd = ev.transmission(test_rho[0])
    np.testing.assert_allclose(fwd, Tt[0], rtol=1e-12)


def test_modal_transmission_is_physical(pkg, gpu, setup):
    """A passive bend transmits at most what enters: T <= 1 at every
    wavelength for every labelled design and for an optimised one."""
    sg, ev, net, norm, history, T, test_rho, Tt, Gt = setup
    assert np.all(T >= 0) and T.max() <= 1 + 1e-3, T.max()
    out = sg.optimize(ev, test_rho[3], steps=12, step=0.05)
    spec = ev.transmission(out["rho"])
    _say(f"plain loop T {out['history'][0]:.4f} -> {out['final']:.4f}")
    assert out["final"] > out["history"][0]
    assert spec.max() <= 1 + 1e-3, spec


def test_network_corrects_the_lagged_gradient(pkg, gpu, setup):
    """What the trained network gains over the method without it: after an
    adjoint at rho_a and a step to rho_b, the anchored gradient g_a + g_net(rho_b)
    - g_net(rho_a) is closer to the adjoint gradient at rho_b than the lagged
    g_a that the loop would use without the network (mean over 8 held-out
    pairs)."""
    sg, ev, net, norm, history, T, test_rho, Tt, Gt = setup
    assert history[-1] < 0.5 * history[0], history
    err_net, err_lag = [], []
    for rho_a, g_a in zip(test_rho[:8], Gt[:8]):
        rho_b = sg._ascent(rho_a, g_a, 0.1)
        _, g_b = ev.adjoint(rho_b)
        corrected = g_a + sg.predict(net, norm, rho_b)[1] - sg.predict(net, norm, rho_a)[1]
        scale = np.linalg.norm(g_b)
        err_net.append(np.linalg.norm(corrected - g_b) / scale)
        err_lag.append(np.linalg.norm(g_a - g_b) / scale)
    _say(f"gradient error at rho_b: network {np.mean(err_net):.3f}, lagged "
         f"{np.mean(err_lag):.3f}")
    assert np.mean(err_net) < np.mean(err_lag)


def test_surrogate_loop_is_verified_and_improves(pkg, gpu, setup):
    """Every accepted design's T is an FDTD value; the loop raises T."""
    sg, ev, net, norm, history, T, test_rho, Tt, Gt = setup
    rho0 = test_rho[1]
    start = ev.transmission(rho0)[sg.objective_index()]
    out = sg.optimize(ev, rho0, steps=8, step=0.02, surrogate=(net, norm))
    assert len(out["history"]) == 9
    final = ev.transmission(out["rho"])[sg.objective_index()]
    # (the final T may come from the adjoint's own forward: same fields, DFT
    # sums replayed in other chunks)
    assert abs(final - out["final"]) <= 1e-12 * abs(final)
    assert final > start


@pytest.fixture(scope="module")
def setup100(tmp_path_factory, pkg):
    """The surrogate's own geometry (200x200x80 bend at 20 nm, 100x100-voxel
    design, 9-wavelength modal planes) on a shortened 1000-step run: 48
    labels and a briefly trained 100x100 FNO with the workload's 9 outputs."""
    sg = pkg.surrogate
    case = sg.surrogate_case(tmp_path_factory.mktemp("surrogate100") / "bend", steps=1000)
    ev = sg.Evaluator(case, checkpoint_every=64)
    rhos = sg.random_designs(48, 100, seed=3)
    _say("labelling 48 designs of the 100x100 bend")
    T, G = ev.label(rhos)
    net = pkg.fno.FNO(100, batch=16, outputs=len(sg.WAVELENGTHS), seed=0)
    norm, history = sg.train(net, rhos, T, G, epochs=30, lr=3e-3)
    return sg, ev, sg.inference_copy(net), norm, history, T


def test_surrogate_loop_100x100(pkg, gpu, setup100):
    """The loop at the workload's design size: one FDTD forward verifies
    every proposal, the final T is the FDTD value of the final design and
    above the start, and the verified spectrum stays below 1 (a passive
    bend)."""
    sg, ev, net, norm, history, T = setup100
    assert history[-1] < history[0]
    assert T.max() <= 1 + 1e-3
    rho0 = sg.random_designs(1, 100, seed=7)[0]
    out = sg.optimize(ev, rho0, steps=6, step=0.01, surrogate=(net, norm))
    assert out["counts"]["forward"] == len(out["errors"]) == 6
    assert out["counts"]["adjoint"] == 1 + out["fallbacks"]
    spec = ev.transmission(out["rho"])
    f = sg.objective_index()
    assert abs(spec[f] - out["final"]) <= 1e-12 * abs(spec[f])
    assert out["final"] > out["history"][0]
    assert spec.max() <= 1 + 1e-3
