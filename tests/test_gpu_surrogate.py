"""The trained FNO surrogate against the method without it (small bend).

A 24x24-voxel bend (40 nm cells, 700 steps) keeps labelling, training and
the optimisation loops within a minute; the workload is the same code as the
100x100 surrogate bench.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIDE = 24


@pytest.fixture(scope="module")
def setup(tmp_path_factory, pkg):
    sg = pkg.surrogate
    case = sg.surrogate_case(tmp_path_factory.mktemp("surrogate") / "bend", side=SIDE,
                             spacing=40e-9, steps=700, pml=6, margin=12, nz=24,
                             pulse_duration=20e-15)
    ev = sg.Evaluator(case, batch=16, checkpoint_every=128)
    train_rho = sg.random_designs(512, SIDE, seed=1)
    test_rho = sg.random_designs(32, SIDE, seed=2)
    T, G = ev.label(train_rho)
    Tt, Gt = ev.label(test_rho)
    net = pkg.fno.FNO(SIDE, batch=16, seed=0)
    norm, history = sg.train(net, train_rho, T, G, epochs=80, lr=3e-3)
    return sg, ev, net, norm, history, test_rho, Tt, Gt


def test_labels_are_consistent(pkg, gpu, setup):
    """A batched label equals the single evaluation of the same design."""
    sg, ev, net, norm, history, test_rho, Tt, Gt = setup
    spec, g = ev.adjoint(test_rho[0])
    np.testing.assert_allclose(spec, Tt[0], rtol=0, atol=0)
    np.testing.assert_array_equal(g, Gt[0])
    fwd = ev.transmission(test_rho[0])
    np.testing.assert_allclose(fwd, Tt[0], rtol=1e-12)


def test_trained_network_beats_the_baseline_predictor(pkg, gpu, setup):
    """On held-out designs the network predicts T better than the training
    mean (the method without it) and its gradient points along the adjoint's."""
    sg, ev, net, norm, history, test_rho, Tt, Gt = setup
    assert history[-1] < 0.5 * history[0], history
    pred = np.array([sg.predict(net, norm, r)[0] for r in test_rho])
    err_net = np.mean((pred - Tt) ** 2)
    err_mean = np.mean((norm.t_mean[None] - Tt) ** 2)
    assert err_net < 0.5 * err_mean, (err_net, err_mean)
    cos = []
    for r, g in zip(test_rho, Gt):
        gh = sg.predict(net, norm, r)[1]
        cos.append(float((gh * g).sum() / (np.linalg.norm(gh) * np.linalg.norm(g))))
    assert np.mean(cos) > 0.3, cos


def test_surrogate_loop_is_verified_and_improves(pkg, gpu, setup):
    """Every accepted design's T is an FDTD value; the loop raises T."""
    sg, ev, net, norm, history, test_rho, Tt, Gt = setup
    rho0 = test_rho[1]
    start = ev.transmission(rho0)[sg.objective_index()]
    out = sg.optimize(ev, rho0, steps=8, step=0.05, surrogate=(net, norm))
    assert len(out["history"]) == 8
    final = ev.transmission(out["rho"])[sg.objective_index()]
    assert final == out["final"]
    assert final > start
