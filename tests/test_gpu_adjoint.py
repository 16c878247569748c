"""Adjoint kernels and driver on the GPU: exact transposes, the Warp gradient
fixture, checkpoint/recompute equivalence and the material pullback."""

import json

import numpy as np
import pytest
import torch

from conftest import BUNDLES, EXPECTED, rel_l2

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def test_cpml_full_state_transpose(pkg, gpu):
    """test_adjoint.py:11-57: <F(x), y> = <x, F^T(y)> for one E+H step with
    CPML psi state on a 6x7x8 grid (seed 71)."""
    k = pkg.kernels
    A = pkg.adjoint
    rng = np.random.default_rng(71)
    shape = (6, 7, 8)
    n = int(np.prod(shape))
    initial = [rng.normal(size=(c, n)).astype(np.float32) * 0.1 for c in (3, 3, 6, 6)]
    cot = [rng.normal(size=a.shape).astype(np.float32) * 0.1 for a in initial]
    e, h, pe, ph = [_t(a) for a in initial]
    qe, qh, qpe, qph = [_t(a) for a in cot]
    eps = _t(rng.uniform(0.2, 1, n))
    ce = np.zeros((3, max(shape), 3), np.float32)
    ce[:, :, 1:] = 1
    ch = ce.copy()
    for axis, length in enumerate(shape):
        ce[axis, [0, length - 1]] = [-0.1, 0.8, 1.2]
        ch[axis, [0, length - 1]] = [-0.13, 0.75, 1.1]
    ce, ch = _t(ce), _t(ch)
    step = torch.zeros(1, dtype=torch.int32, device="cuda")
    dummy = torch.zeros((3, n), device="cuda")
    shared = [dummy, dummy, step, dummy, 1, *shape, -1, 0.4, 0]
    k.launch(k.update_e, n, inputs=[e, h, eps, pe, ce, *shared])
    k.launch(k.update_h, n, inputs=[h, e, eps, ph, ch, *shared])
    outputs = [a.cpu().numpy() for a in (e, h, pe, ph)]
    d = torch.empty_like(pe)
    hist = torch.zeros((1, 3, 1), device="cuda")
    grad = torch.zeros(1, device="cuda")
    k.launch(A.prepare, n, inputs=[qh, eps, qph, ch, d, *shape, 0.4, 0])
    args = [d, hist, grad, step, *shape, 0, 0, 0, 0, 0, 0]
    k.launch(A.gather, n, inputs=[qe, *args, 1])
    k.launch(A.prepare, n, inputs=[qe, eps, qpe, ce, d, *shape, 0.4, 1])
    k.launch(A.gather, n, inputs=[qh, *args, 0])
    transposed = [a.cpu().numpy() for a in (qe, qh, qpe, qph)]
    lhs = sum(np.sum(a.astype(float) * b) for a, b in zip(outputs, cot))
    rhs = sum(np.sum(a.astype(float) * b) for a, b in zip(initial, transposed))
    assert abs(lhs - rhs) / max(abs(lhs), abs(rhs)) < 2e-6


def test_space_time_colocation_transpose(pkg, gpu):
    """test_adjoint.py:60-127: detector record (exact colocation, phasor)
    and adjoint inject are transposes (seed 17)."""
    k = pkg.kernels
    D = pkg.detectors
    A = pkg.adjoint
    rng = np.random.default_rng(17)
    shape = (6, 7, 8)
    n = int(np.prod(shape))
    e, h, hp = [_t(rng.normal(size=(3, n))) for _ in range(3)]
    sx, sy, sz = 1, 7, 8
    count = sx * sy * sz
    step = torch.zeros(1, dtype=torch.int32, device="cuda")
    indices = torch.zeros(1, dtype=torch.int32, device="cuda")
    weights = _t(np.array([[0.3, 0.7]]))
    ph = torch.zeros((6, count, 2), device="cuda")
    flux = torch.zeros(1, device="cuda")
    area = torch.ones(count, device="cuda")
    k.launch(D.record, count, inputs=[e, h, hp, step, weights, indices, ph, flux, area, *shape,
                                      2, 0, 0, sy, sz, 0, 1, 1])
    source_np = rng.normal(size=(6, count, 2)).astype(np.float32)
    source = _t(source_np)
    qe = torch.zeros_like(e)
    qh = torch.zeros_like(h)
    args = [qe, qh, source, weights, step, *shape, 2, 0, 0, sy, sz]
    k.launch(A.inject, count, inputs=[*args, 0])
    rhs = np.sum(qe.cpu().numpy().astype(float) * e.cpu().numpy()) + np.sum(
        qh.cpu().numpy().astype(float) * h.cpu().numpy())
    qh.zero_()
    k.launch(A.inject, count, inputs=[*args, 1])
    rhs += np.sum(qh.cpu().numpy().astype(float) * hp.cpu().numpy())
    v = ph.cpu().numpy()
    lhs = np.sum(source_np[..., 0].astype(float) * v[..., 0]
                 - source_np[..., 1].astype(float) * v[..., 1])
    assert abs(lhs - rhs) / max(abs(lhs), abs(rhs)) < 2e-6


def test_material_forward_matches_bundle(pkg, gpu):
    """gradient_benchmark.py:27-31: Material(p0, beta) reproduces the bundle's
    design-plane inv_permittivity (reference max abs 1.6e-6)."""
    data = np.load(BUNDLES / "gradient20_highest_exact/inputs.npz")
    meta = json.loads((BUNDLES / "gradient20_highest_exact/manifest.json").read_text())
    p0 = data["parameters/p0"].reshape(80, 80)
    mat = pkg.material.Material(p0, meta["beta"])
    want = data["inv_permittivity"][0, 100:180, 20:100, 30]
    assert np.max(np.abs(mat.inv_eps.numpy() - want)) < 2e-6


def test_material_pullback_finite_difference(pkg, gpu):
    """test_material.py:7-22 (seed 51): pullback vs central FD < 1%."""
    Material = pkg.material.Material
    rng = np.random.default_rng(51)
    p = rng.uniform(0.2, 0.8, (16, 16)).astype(np.float32)
    weights = rng.normal(size=p.shape).astype(np.float32)
    direction = rng.normal(size=p.shape).astype(np.float32)
    direction /= np.linalg.norm(direction)
    gradient = Material(p, 50).pullback(weights)
    analytic = float(np.sum(gradient.astype(float) * direction))
    errors = []
    for hstep in (0.01, 0.003, 0.001):
        plus = Material(p + hstep * direction, 50).inv_eps.numpy()
        minus = Material(p - hstep * direction, 50).inv_eps.numpy()
        finite = float(np.sum((plus.astype(float) - minus) * weights) / (2 * hstep))
        errors.append(abs(finite - analytic) / max(abs(analytic), 1e-12))
    assert min(errors) < 0.01


@pytest.fixture(scope="module")
def gradient20(pkg, gpu):
    data = np.load(BUNDLES / "gradient20_highest_exact/inputs.npz")
    p0 = data["parameters/p0"].reshape(80, 80)
    adj = pkg.adjoint.Adjoint(BUNDLES / "gradient20_highest_exact")
    pkg.detectors.attach(adj)
    loss, gradient = adj.differentiate(p0)
    voxel = adj.voxel_gradient()
    meas = pkg.detectors.measurements(adj)
    del adj
    torch.cuda.empty_cache()
    return p0, loss, gradient, voxel, meas


def test_gradient20_matches_warp_fixture(gradient20):
    """3M-cell bend, 5,245 steps, full curl history: transmission, voxel
    gradient (input of the pullback) and parameter gradient against Warp's
    results/warp/gradient20 (Warp vs FDTDX exact: 1.85e-6)."""
    p0, loss, gradient, voxel, meas = gradient20
    ref = json.loads((EXPECTED / "gradient20/result.json").read_text())
    want = np.load(EXPECTED / "gradient20/warp_gradient.npz")
    assert abs(meas["modal_transmission"] - ref["measurements"]["modal_transmission"]) < 2e-6
    assert abs(-loss - ref["measurements"]["modal_transmission"]) < 2e-6
    assert rel_l2(voxel, want["voxel_gradient"]) < 1e-4
    assert rel_l2(gradient, want["gradient"]) < 1e-4


def test_checkpointed_adjoint_equals_full_history(pkg, gradient20):
    """Checkpoints every 64 steps + recompute give the full-history voxel
    gradient bit for bit (the forward has no atomics)."""
    p0, loss, gradient, voxel, meas = gradient20
    adj = pkg.adjoint.Adjoint(BUNDLES / "gradient20_highest_exact", checkpoint_every=64)
    pkg.detectors.attach(adj)
    loss2, gradient2 = adj.differentiate(p0)
    assert loss2 == loss
    np.testing.assert_array_equal(adj.voxel_gradient(), voxel)
    np.testing.assert_array_equal(gradient2, gradient)
    assert adj.history.shape[0] == 64


def test_odd_checkpoint_chunks_equal_full_history(pkg, gradient20):
    """Checkpoints every 321 steps (chunks of 321 and a 109-step tail: every
    reverse chunk odd, so each ends by copying the second psi buffers back)
    give the full-history gradient bit for bit."""
    p0, loss, gradient, voxel, meas = gradient20
    adj = pkg.adjoint.Adjoint(BUNDLES / "gradient20_highest_exact", checkpoint_every=321)
    pkg.detectors.attach(adj)
    loss2, gradient2 = adj.differentiate(p0)
    assert loss2 == loss
    np.testing.assert_array_equal(gradient2, gradient)


def test_turned_inject_equals_the_reference_launch_order(pkg, gradient20):
    """Captured reverse steps ending with wfd_adjoint_inject_turn (inject(1),
    retreat and the next step's inject(0) in one launch) give the gradient
    of the reference's five launches per step bit for bit."""
    p0, loss, gradient, voxel, meas = gradient20
    out = []
    for turned in (True, False):
        adj = pkg.adjoint.Adjoint(BUNDLES / "gradient20_highest_exact", checkpoint_every=64)
        adj.turned_inject = turned
        pkg.detectors.attach(adj)
        loss2, gradient2 = adj.differentiate(p0)
        assert adj._turns() == turned
        assert int(adj.reverse_step.numpy()[0]) == -1
        out.append((loss2, adj.voxel_gradient()))
        adj.close()
    assert out[0][0] == out[1][0] == loss
    np.testing.assert_array_equal(out[0][1], out[1][1])
    np.testing.assert_array_equal(out[0][1], voxel)


@pytest.mark.parametrize("from_h", [1, 0])
def test_fused_reverse_half_equals_prepare_then_gather(pkg, gpu, from_h):
    """wfd_adjoint_reverse_half (prepare + gather in one pass, psi to a second
    buffer) equals prepare followed by gather bit for bit: q, psi and the FP64
    gradient, with CPML slabs on every face and a design box (seed 5)."""
    A = pkg.adjoint
    rng = np.random.default_rng(5)
    shape = (11, 9, 13)
    n = int(np.prod(shape))
    box = (2, 1, 3, 6, 5, 7)
    steps = 4
    q_src = _t(rng.normal(size=(3, n)))
    q_dst = _t(rng.normal(size=(3, n)))
    psi = rng.normal(size=(6, n)).astype(np.float32)
    eps = _t(rng.uniform(0.1, 1, n))
    coeff = np.zeros((3, max(shape), 3), np.float32)
    coeff[:, :, 2] = 1
    for axis, length in enumerate(shape):
        for p in (0, 1, length - 2, length - 1):
            coeff[axis, p] = rng.uniform(-0.5, 0.9, 3)
    coeff = _t(coeff)
    hist = _t(rng.normal(size=(steps, 3, box[3] * box[4] * box[5])))
    step = torch.tensor([2], dtype=torch.int32, device="cuda")
    grad0 = torch.from_numpy(rng.normal(size=box[3] * box[4] * box[5])).cuda()

    # two-kernel reference
    qa, psia, ga = q_dst.clone(), _t(psi), grad0.clone()
    d = torch.empty((6, n), device="cuda")
    A._prepare(q_src, eps, psia, coeff, d, *shape, 0.57, 1 - from_h)
    A._gather(qa, d, hist, ga, step, *shape, *box, from_h)
    # fused, psi_out starting as psi_in (slots without CPML are not written)
    qb, gb = q_dst.clone(), grad0.clone()
    psi_in, psi_out = _t(psi), _t(psi)
    A._reverse_half(qb, q_src, eps, psi_in, psi_out, coeff, hist, gb, step, *shape, *box,
                    0.57, from_h)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(qb.cpu().numpy(), qa.cpu().numpy())
    np.testing.assert_array_equal(psi_out.cpu().numpy(), psia.cpu().numpy())
    np.testing.assert_array_equal(psi_in.cpu().numpy(), psi)  # read only
    np.testing.assert_array_equal(gb.cpu().numpy(), ga.cpu().numpy())
    assert not np.array_equal(qb.cpu().numpy(), q_dst.cpu().numpy())


@pytest.mark.parametrize("overlap", [False, True])
def test_inject_many_equals_one_launch_per_plane(pkg, gpu, overlap):
    """wfd_adjoint_inject_many (every plane in one launch; a target in several
    planes' footprints is owned by the first plane's thread, which adds them
    in order) equals wfd_adjoint_inject per plane in order, bit for bit, for
    E+H and H_old-only injections (seed 7)."""
    A = pkg.adjoint
    lib = pkg._lib
    rng = np.random.default_rng(7)
    shape = (20, 12, 14)
    n = int(np.prod(shape))
    steps = 5
    # (x0, y0, z0, sx, sy, sz): two x-normal planes and a third far away, or
    # a y-normal plane crossing both footprints (as the two mode planes of a
    # bend do) followed by an x-normal plane next to the first
    planes = [(3, 2, 1, 1, 8, 10), (12, 1, 2, 1, 9, 9), (17, 3, 0, 1, 7, 11)]
    if overlap:
        planes[2:] = [(2, 5, 0, 12, 1, 11), (4, 3, 0, 1, 7, 11)]
    srcs = []
    for x0, y0, z0, sx, sy, sz in planes:
        count = sx * sy * sz
        source = _t(rng.normal(size=(6, count, 2)))
        weights = rng.normal(size=(steps, 2)).astype(np.float32)
        weights[1] = 0  # a step without injection
        srcs.append((source, _t(weights), (x0, y0, z0), (sy, sz), count))
    table = (lib.InjectSource * len(srcs))(*[
        lib.InjectSource(s.data_ptr(), w.data_ptr(), *lo, *syz, c, steps)
        for s, w, lo, syz, c in srcs])
    qe0 = _t(rng.normal(size=(3, n)))
    qh0 = _t(rng.normal(size=(3, n)))
    for t in (0, 1, 4):
        step = torch.tensor([t], dtype=torch.int32, device="cuda")
        for old in (0, 1):
            qa, qha = qe0.clone(), qh0.clone()
            for s, w, lo, syz, c in srcs:
                A._inject(qa, qha, s, w, step, *shape, *lo, *syz, old, dim=c)
            qb, qhb = qe0.clone(), qh0.clone()
            A._inject_many(table, qb, qhb, step, *shape, old)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(qb.cpu().numpy(), qa.cpu().numpy())
            np.testing.assert_array_equal(qhb.cpu().numpy(), qha.cpu().numpy())
            changed = not np.array_equal(qhb.cpu().numpy(), qh0.cpu().numpy())
            assert changed == (t != 1)
