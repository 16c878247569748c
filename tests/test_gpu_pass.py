"""The single-pass step (wfd_yee_pass: one launch per step, the state
alternating between two buffer sets) against the fused kernel pair
(wfd_yee_step, itself bitwise equal to the CPU oracle) and the oracle.

It performs the fused step's arithmetic operation for operation, so fields,
CPML psi, counters and every staged detector must agree bit for bit over
whole runs, one and two steps per launch, chunk lengths that end on
one-step passes and a fused step, every launch shape, and through the
adjoint's checkpointed forward and recompute.
"""

import sys

import numpy as np
import pytest
import torch

from conftest import BUNDLES, rel_l2
from oracle.model import OracleSolver

pytestmark = pytest.mark.gpu


def _say(text):
    print(f"[pass] {text}", file=sys.stderr, flush=True)


def _solver(pkg, case, single_pass, attach=True, **kw):
    s = pkg.solver.Solver(case, resident=False, single_pass=single_pass, **kw)
    if attach:
        pkg.detectors.attach(s)
    s.compile()
    s.reset()
    assert s.single_pass == (single_pass is True)
    return s


def _assert_same(a, b):
    torch.cuda.synchronize()
    for name in ("E", "H", "psi_e", "psi_h", "counters"):
        np.testing.assert_array_equal(getattr(a, name).numpy(), getattr(b, name).numpy(),
                                      err_msg=name)
    for ma, mb in zip(a.monitors, b.monitors):
        for attr in ("phasor", "flux"):
            va, vb = getattr(ma, attr, None), getattr(mb, attr, None)
            if hasattr(va, "numpy"):
                np.testing.assert_array_equal(va.numpy(), vb.numpy(),
                                              err_msg=f"{ma.name}.{attr}")


@pytest.fixture(scope="module")
def bend(tmp_path_factory, pkg):
    """The 200x200x80 bend at 20 nm (3.2M cells, past the resident capacity)
    with modal DFT planes at 9 wavelengths, 300 steps."""
    path = tmp_path_factory.mktemp("pass") / "bend"
    pkg.surrogate.surrogate_case(path, steps=300)
    return path


@pytest.mark.parametrize("steps", [1, 2])
def test_forward50_run_bitwise(pkg, gpu, steps):
    """The whole 2,098-step forward50 run (80x80x30, two staged flux planes)
    with one or two steps per launch."""
    a = _solver(pkg, BUNDLES / "forward50_gpu", False)
    b = _solver(pkg, BUNDLES / "forward50_gpu", True)
    assert len(b.staged()) == len(b.monitors) == 2
    b.use_single_pass(0, 0, steps)
    b.capture_all()
    assert b.single_pass and b.pass_shape[2] == steps
    a.run()
    b.run()
    _assert_same(a, b)
    assert int(b.step.numpy()[0]) == b.steps


@pytest.mark.parametrize("shape", [(4, 2, 1), (16, 4, 1), (5, 4, 1), (2, 1, 1), (1, 1, 1),
                                   (3, 1, 1, 1024), (8, 2, 2), (4, 1, 2), (13, 2, 2),
                                   (1, 2, 2), (5, 1, 2, 1024)])
def test_launch_shapes_and_odd_chunks_bitwise(pkg, gpu, bend, shape):
    """Planes per workgroup x slots per thread x steps per launch (x threads
    per workgroup), each over ragged run_steps (lengths of 4k + 1..3 end with
    one-step passes and one fused step) on the 200x200x80 bend."""
    a = _solver(pkg, bend, False)
    b = _solver(pkg, bend, True)
    b.use_single_pass(*shape)
    b.capture_all()
    for count in (37, 64, 1, 70, 3):
        a.run_steps(count)
        b.run_steps(count)
        _assert_same(a, b)
    a.close()
    b.close()


def test_bend_against_the_oracle(pkg, gpu, bend):
    """100 steps of the bend on the single pass against the CPU oracle: fields
    and psi bitwise, the FP64 DFT planes within 1e-12."""
    s = _solver(pkg, bend, True)
    s.run_steps(100)
    _say("GPU 100 steps done; oracle")
    o = OracleSolver(bend)
    o.attach()
    o.run(100)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(s.E.numpy(), o.E)
    np.testing.assert_array_equal(s.H.numpy(), o.H)
    np.testing.assert_array_equal(s.psi_e.numpy(), o.psi["E"])
    np.testing.assert_array_equal(s.psi_h.numpy(), o.psi["H"])
    for m, om in zip(s.monitors, o.monitors):
        assert m.name == om.name
        assert rel_l2(m.phasor.numpy(), om.phasor) < 1e-12
    s.close()


@pytest.mark.parametrize("steps,checkpoint", [(1, 64), (2, 64), (2, None)])
def test_checkpointed_adjoint_on_the_single_pass_bitwise(pkg, gpu, bend, steps, checkpoint):
    """The adjoint's forward, checkpoints and recompute on one- or two-step
    passes (one-step chunks: record_curl reads H from whichever set holds
    the step's start; two-step passes write the curl history of both their
    steps themselves), with checkpoints every 64 steps or the full history:
    T and dT/drho bitwise those of the fused pair."""
    rho = pkg.surrogate.random_designs(1, 100, seed=5)[0]
    out = []
    for single in (False, True):
        p = pkg.design.TransmissionProblem(bend, wavelength=pkg.surrogate.OBJECTIVE,
                                           checkpoint_every=checkpoint, single_pass=single)
        p._compile()
        assert p.adjoint.single_pass == single
        if single:
            p.adjoint.use_single_pass(0, 0, steps)
            p.adjoint.capture_all()
        T, g = p.value_and_grad(rho)
        out.append((T, g.cpu().numpy()))
        p.adjoint.close()
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[0][1], out[1][1])
    assert np.abs(out[1][1]).max() > 0


@pytest.mark.parametrize("steps", [0, 1, 2])
def test_monitors_on_one_footprint_share_a_stage(pkg, gpu, tmp_path, steps):
    """A bend with modal and flux DFT planes on the same two planes: the four
    staged monitors share two stage buffers, so the step stages each plane
    once.  Their phasors equal bit for bit those of runs that stage the modal
    planes alone and the flux planes alone (no sharing), on the fused pair
    (steps 0) and on one- and two-step passes, and agree with per-step
    recording (packed CPML storage, no staging; its float64 sums differ in
    the last bits) to 1e-12."""
    rho = np.random.default_rng(3).uniform(0, 1, (12, 12))
    case = pkg.cases.bend_case(tmp_path / "bend", rho, shape=(48, 48, 24), spacing=40e-9,
                               steps=150, pml=6, design=(12, 12), pulse_duration=20e-15,
                               flux_wavelengths=[1.5e-6, 1.6e-6])

    def run(keep=None):
        s = pkg.solver.Solver(case, resident=False, single_pass=steps > 0)
        pkg.detectors.attach(s)
        if keep is not None:
            s.monitors = [m for m in s.monitors if m.name.startswith(keep)]
        s.compile()
        if steps:
            s.use_single_pass(0, 0, steps)
            s.capture_all()
        s.reset()
        s.run()
        torch.cuda.synchronize()
        return s

    shared = run()
    names = [m.name for m in shared.staged()]
    assert names == ["mode_in", "mode_out", "flux_in", "flux_out"]
    assert shared.monitors[0].stage is shared.monitors[2].stage
    assert shared.monitors[1].stage is shared.monitors[3].stage
    assert shared.monitors[0].stage is not shared.monitors[1].stage
    assert shared._plan().nstage == 2
    alone = {m.name: m for keep in ("mode", "flux") for m in run(keep).monitors}
    assert sorted(alone) == sorted(names)
    for m in shared.monitors:
        np.testing.assert_array_equal(m.phasor.numpy(), alone[m.name].phasor.numpy(),
                                      err_msg=m.name)
    plain = pkg.solver.Solver(case, cpml_storage="packed")
    pkg.detectors.attach(plain)
    plain.compile()
    plain.reset()
    plain.run()
    assert plain.staged() == []
    for a, b in zip(plain.monitors, shared.monitors):
        assert a.name == b.name
        assert rel_l2(b.phasor.numpy(), a.phasor.numpy()) < 1e-12, a.name
    np.testing.assert_array_equal(plain.E.numpy(), shared.E.numpy())
