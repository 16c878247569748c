"""Batched stepping of uneven grids (wfd_batch_step) against single-grid runs."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

STEPS = 300


@pytest.fixture(scope="module")
def cases(tmp_path_factory, pkg):
    root = tmp_path_factory.mktemp("batch")
    c = pkg.cases
    rho = np.random.default_rng(1).uniform(0, 1, (12, 12))
    return [
        c.bend_case(root / "bend", rho, shape=(48, 48, 24), spacing=40e-9, steps=STEPS, pml=6,
                    design=(12, 12), pulse_duration=20e-15),
        c.waveguide_case(root / "guide", shape=(64, 40, 32), steps=STEPS, pml=8),
        c.slab_case(root / "slab", n=24, spacing=50e-9, steps=STEPS, pml=4, thickness=6),
    ]


def _solvers(pkg, cases, chunks):
    out = []
    for case, xc in zip(cases, chunks):
        s = pkg.solver.Solver(case, x_chunk=xc, resident=False)
        pkg.detectors.attach(s)
        out.append(s)
    return out


@pytest.mark.parametrize("chunks,single_pass", [((1, 2, 4), False), ((2, 2, 8), False),
                                                ((1, 2, 4), "batched"), ((1, 2, 4), "each")])
def test_batch_equals_single_runs(pkg, gpu, cases, chunks, single_pass):
    """Three shapes, mixed x-chunk classes, staged DFT planes: every design's
    fields, CPML state, counters and phasors equal its own solo run bitwise,
    on the batched fused step or on the batched two-step pass (300 steps:
    chunks of 64 and 44, both multiples of 4, all on passes)."""
    solo = _solvers(pkg, cases, chunks)
    for s in solo:
        s.compile()
        s.reset()
        s.run()
    batch = pkg.batch.Batch(_solvers(pkg, cases, chunks), single_pass=single_pass)
    batch.compile()
    assert batch.use_pass == bool(single_pass)
    assert not single_pass or batch.pass_each == (single_pass == "each")
    batch.run()
    torch.cuda.synchronize()
    for a, b in zip(solo, batch.solvers):
        for name in ("E", "H", "psi_e", "psi_h", "counters"):
            np.testing.assert_array_equal(getattr(a, name).numpy(), getattr(b, name).numpy(),
                                          err_msg=name)
        for ma, mb in zip(a.monitors, b.monitors):
            np.testing.assert_array_equal(ma.phasor.numpy(), mb.phasor.numpy())
    # replay after reset gives the same bits again
    first = [s.E.numpy().copy() for s in batch.solvers]
    batch.reset()
    batch.run()
    for f, s in zip(first, batch.solvers):
        np.testing.assert_array_equal(f, s.E.numpy())


def test_batch_validation(pkg, gpu, cases):
    s = _solvers(pkg, cases, (2, 2, 2))
    with pytest.raises(ValueError):
        pkg.batch.Batch([])
    packed = pkg.solver.Solver(cases[0], cpml_storage="packed")
    with pytest.raises(ValueError):
        pkg.batch.Batch([s[0], packed])
    with pytest.raises(ValueError):
        pkg.batch.Batch(s, graph_steps=32)


@pytest.mark.parametrize("checkpoint,single_pass", [(128, "auto"), (None, "batched"),
                                                   (None, False)])
def test_batched_forward_adjoint_equals_single_problems(pkg, gpu, tmp_path, checkpoint,
                                                        single_pass):
    """BatchTransmission (batched forwards with checkpoints between chunk
    graphs, or recording every design's whole curl history on the batched
    two-step pass or before each fused step; per-design reverses) equals
    each design's own TransmissionProblem bit for bit: T and dT/drho."""
    rng = np.random.default_rng(7)
    cases, rhos = [], []
    for k, (shape, side) in enumerate((((48, 48, 24), 12), ((56, 48, 24), 14),
                                       ((48, 56, 32), 12))):
        rho = pkg.cases.design_density(rng, (side, side), blurred=bool(k % 2))
        cases.append(pkg.cases.bend_case(tmp_path / f"b{k}", rho, shape=shape, spacing=40e-9,
                                         steps=STEPS, pml=6, design=(side, side),
                                         pulse_duration=20e-15,
                                         flux_wavelengths=[1.5e-6, 1.55e-6, 1.6e-6]))
        rhos.append(rng.uniform(0, 1, (side, side)))
    bt = pkg.design.BatchTransmission(cases, wavelength=1.55e-6, checkpoint_every=checkpoint,
                                      single_pass=single_pass)
    T = bt.forward(rhos)
    if single_pass == "batched":
        assert bt.batch.use_pass and not bt.batch.pass_each
    elif single_pass is False:
        assert not bt.batch.use_pass
    grads = bt.backward()
    for case, rho, t, g in zip(cases, rhos, T, grads):
        p = pkg.design.TransmissionProblem(case, wavelength=1.55e-6, checkpoint_every=checkpoint)
        t1, g1 = p.value_and_grad(rho)
        assert t == t1
        np.testing.assert_array_equal(g.cpu().numpy(), g1.cpu().numpy())
    with pytest.raises(ValueError):
        pkg.design.BatchTransmission(cases, checkpoint_every=100)
