"""DrivenUniform, periodic and volume paths on the GPU against the oracle and
the reference's fixtures/closed forms."""

import json

import numpy as np
import pytest
import torch

from conftest import EXPECTED, rel_l2
from oracle import kernels as OK

pytestmark = pytest.mark.gpu


def _oracle_driven(d, steps):
    """The oracle's flat restatement stepped with the solver's own tables."""
    g = d._desc
    e = np.zeros(len(d.e), np.float32)
    h = np.zeros(len(d.h), np.float32)
    wave = d.wave.numpy()
    counter = np.zeros(1, np.int32)
    n = list(d.shape)
    L = [g.length, g.width, g.height]
    for _ in range(steps):
        OK.driven_magnetic(e, h, n, g.spacing, g.dt, L, g.sponge_width, g.sponge_rate)
        OK.driven_electric(h, e, n, g.spacing, g.dt, L, [g.core_width, g.core_height],
                           g.core_epsilon, g.clad_epsilon, g.sponge_width, g.sponge_rate,
                           [g.source_x, g.source_y, g.source_z],
                           [g.source_sigma_x, g.source_sigma_y, g.source_sigma_z], wave,
                           counter, g.normal_average)
        counter[0] += 1
    return e, h


@pytest.mark.parametrize("layout", ["flat", "xyz"])
def test_driven_uniform_matches_oracle(pkg, gpu, layout):
    """driven_uniform.py kernels vs the oracle over 64 steps (exp() of the
    impressed current may differ by an ulp between libm and the GPU)."""
    from warp_fdtd_amd.driven_uniform import DrivenUniform
    from warp_fdtd_amd.guide_problem import GuideProblem

    p = GuideProblem(length=0.8, width=0.8, height=0.8, source_x=0.4, pulse_center=0.3,
                     pulse_sigma=0.1, sponge_width=0.1)
    d = DrivenUniform(p, 0.1, 128, layout=layout)
    d.compile()
    d.run(64)
    e, h = _oracle_driven(d, 64)
    assert rel_l2(d.e.numpy(), e) < 1e-5
    assert rel_l2(d.h.numpy(), h) < 1e-5
    assert int(d.counter.numpy()[0]) == 64


def test_layouts_observer_passivity_and_counter(pkg, gpu):
    """test_driven_guide.py:94-111 and test_driven_cuda.py:13-53: flat == xyz,
    the observer is passive and repeatable, graph + eager tail counter."""
    from warp_fdtd_amd.driven_uniform import DrivenUniform
    from warp_fdtd_amd.guide_observer import GuideObserver
    from warp_fdtd_amd.guide_problem import GuideProblem

    problem = GuideProblem()
    a = DrivenUniform(problem, 0.2, 64, layout="flat")
    b = DrivenUniform(problem, 0.2, 64, layout="xyz")
    for s in (a, b):
        s.compile()
        s.run(32)
    np.testing.assert_allclose(a.e.numpy(), b.e.numpy(), rtol=1e-5, atol=1e-8)
    np.testing.assert_allclose(a.h.numpy(), b.h.numpy(), rtol=1e-5, atol=1e-8)
    old_e, old_h = b.e.numpy().copy(), b.h.numpy().copy()
    observer = GuideObserver(b, problem)
    first = observer.sample()
    second = observer.sample()
    np.testing.assert_array_equal(first, second)
    np.testing.assert_array_equal(old_e, b.e.numpy())
    np.testing.assert_array_equal(old_h, b.h.numpy())
    assert int(b.counter.numpy()[0]) == 32
    b.reset()
    b.run(45)
    assert int(b.counter.numpy()[0]) == 45
    before = b.e.numpy()
    b.reset()
    b.run(45)
    np.testing.assert_array_equal(before, b.e.numpy())


def test_u050_observer_samples_match_warp(pkg, gpu):
    """results/octree_guide_controls/u050_small: 393k cells, 462 steps,
    observer samples at 3/7, 5/7 and 7/7 of the run (Warp float32, xyz)."""
    from warp_fdtd_amd.driven_uniform import DrivenUniform
    from warp_fdtd_amd.guide_observer import GuideObserver
    from warp_fdtd_amd.guide_problem import GuideProblem

    meta = json.loads((EXPECTED / "u050_small/result.json").read_text())
    want = np.load(EXPECTED / "u050_small/fields.npz")["fields"]
    p = GuideProblem(**meta["problem"])
    steps = meta["steps"]
    cfl = (p.duration / steps) / (meta["spacing"] * np.sqrt(p.clad_epsilon / 3))
    s = DrivenUniform(p, meta["spacing"], steps, cfl)
    assert abs(s.dt - meta["dt"]) < 1e-12
    s.compile()
    observer = GuideObserver(s, p)
    s.reset()
    previous = 0
    got = []
    for checkpoint in (3 * steps // 7, 5 * steps // 7, steps):
        s.run(checkpoint - previous)
        got.append(observer.sample())
        previous = checkpoint
    for a, b in zip(got, want):
        assert rel_l2(a, b) < 1e-4


def test_nonfinite_counter(pkg, gpu):
    """test_driven_guide.py:113-121: two injected non-finite values."""
    from warp_fdtd_amd.guide_validation import count_nonfinite
    from warp_fdtd_amd.kernels import launch

    values = torch.tensor([1.0, float("nan"), float("inf"), -2.0], device="cuda")
    count = torch.zeros(1, dtype=torch.int32, device="cuda")
    launch(count_nonfinite, 4, inputs=[values, count])
    assert int(count.item()) == 2


def test_assert_finite_on_a_stepped_guide(pkg, gpu):
    """guide_validation.py:12-20: passes on a stepped DrivenUniform; one
    non-finite value in e or h raises FloatingPointError naming the count."""
    from warp_fdtd_amd.driven_uniform import DrivenUniform
    from warp_fdtd_amd.guide_problem import GuideProblem
    from warp_fdtd_amd.guide_validation import assert_finite

    s = DrivenUniform(GuideProblem(), 0.2, 16, layout="flat")
    s.compile()
    s.run(8)
    assert_finite(s)
    s.h.tensor[3] = float("nan")
    s.e.tensor[5] = float("inf")
    with pytest.raises(FloatingPointError, match="2 nonfinite"):
        assert_finite(s)


def test_periodic_dispersion_and_gauss_law(pkg, gpu):
    """test_periodic.py:10-46 (Yee branch): oblique periodic mode frequency
    equals 2/dt asin(S dt/2) within 2e-6 and the discrete divergence stays
    below 2e-5."""
    from warp_fdtd_amd.convergence import PERIOD, fit_frequency
    from warp_fdtd_amd.periodic import Periodic

    n = 16
    x, y, z = np.indices((n, n, n))
    s = Periodic(n, 0.5, "yee", shape=(n, n, n))
    s.initial.fill(0)
    wave = np.cos(2 * np.pi * (x + y + z + 0.5) / n).astype(np.float32).ravel()
    s.initial[0] = wave
    s.initial[1] = -wave
    s.reset()
    s.compile()
    s.run()
    ratio, residual, _ = fit_frequency(s.signal.numpy(), s.dt * np.sqrt(3))
    symbol = np.sin(np.pi / n)
    space = 2 * 299792458 / (1550e-9 / n) * np.sqrt(3) * symbol
    omega = 2 * np.pi / PERIOD * np.sqrt(3)
    pred = 2 / s.dt * np.arcsin(space * s.dt / 2) / omega
    assert abs(ratio / pred - 1) < 2e-6
    assert residual < 1e-4
    fields = s.state.numpy().reshape(6, n, n, n)
    div = sum(fields[c] - np.roll(fields[c], 1, axis=c) for c in range(3))
    assert np.linalg.norm(div) / np.linalg.norm(fields[:3]) < 2e-5


def test_volume_sampler_affine_and_time_centering(pkg, gpu):
    """test_volume.py:13-61: trilinear, time-centred sampling is exact on
    affine staggered fields."""
    from warp_fdtd_amd.kernels import launch
    from warp_fdtd_amd.volume import record_volume

    shape = (7, 8, 9)
    xyz = np.indices(shape).astype(float)
    offsets = np.array([[0.5, 0, 0], [0, 0.5, 0], [0, 0, 0.5], [0, 0.5, 0.5], [0.5, 0, 0.5],
                        [0.5, 0.5, 0]])
    points = np.array([[2.1, 3.4, 4.2], [1.3, 2.5, 3.7], [4.8, 5.1, 6.3]], "f")
    fields = []
    for c in range(6):
        position = xyz + offsets[c, :, None, None, None]
        fields.append((position[0] + 2 * position[1] - 3 * position[2] + c).astype("f").ravel())
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()
    e = dev(np.array(fields[:3]))
    h = dev(4 * np.array(fields[3:]))
    prev = dev(2 * np.array(fields[3:]))
    weights = dev(np.array([[0, 0], [0.2, 0.3]]))
    step = torch.zeros(1, dtype=torch.int32, device="cuda")
    out = torch.zeros((6, len(points), 2), device="cuda")
    args = [e, h, prev, dev(points), weights, step, out, *shape, 1]
    launch(record_volume, len(points), inputs=args)
    assert float(out.abs().max()) == 0
    step.fill_(1)
    launch(record_volume, len(points), inputs=args)
    raw = out.cpu().numpy()
    actual = raw[..., 0] + 1j * raw[..., 1]
    for c in range(6):
        v = points[:, 0] + 2 * points[:, 1] - 3 * points[:, 2] + c
        expected = v * (1 if c < 3 else 3) * (0.2 + 0.3j)
        np.testing.assert_allclose(actual[c], expected, rtol=2e-6, atol=3e-6)


def test_volume_matches_plane_colocation(pkg, gpu):
    """test_volume.py:63-103: at E_z points the volume sampler equals the
    detector's exact colocation, zero halos included (seed 18)."""
    from warp_fdtd_amd.detectors import record
    from warp_fdtd_amd.kernels import launch
    from warp_fdtd_amd.volume import record_volume

    shape = (7, 8, 9)
    n = int(np.prod(shape))
    rng = np.random.default_rng(18)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()
    e, h, prev = [dev(rng.normal(size=(3, n))) for _ in range(3)]
    weights = dev(np.array([[0.3, -0.2]]))
    step = torch.zeros(1, dtype=torch.int32, device="cuda")
    indices = torch.zeros(1, dtype=torch.int32, device="cuda")
    a = torch.zeros((6, n, 2), device="cuda")
    b = torch.zeros_like(a)
    points = np.indices(shape).astype("f")
    points[2] += 0.5
    points = dev(points.reshape(3, -1).T)
    launch(record_volume, n, inputs=[e, h, prev, points, weights, step, a, *shape, 0])
    launch(record, n, inputs=[e, h, prev, step, weights, indices, b, torch.zeros(1, device="cuda"),
                              torch.ones(n, device="cuda"), *shape, 0, 0, 0, shape[1], shape[2],
                              2, 1, 1])
    np.testing.assert_allclose(a.cpu().numpy(), b.cpu().numpy(), rtol=2e-6, atol=2e-7)
