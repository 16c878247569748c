"""Density design map, DFT-plane adjoint and the autograd wrapper on the GPU.

The small bend (48x48x24 at 40 nm, 12x12 design voxels through the 6-cell
core layer, 700 steps) keeps the FP64 CPU oracle fast enough for 32 central
finite differences; it runs the same code paths as the 200x200x80 bend.
"""

import numpy as np
import pytest
import torch

from oracle.model import OracleSolver

pytestmark = pytest.mark.gpu

EPS_LO, EPS_HI = 2.085, 12.08
SHAPE, DESIGN, STEPS = (48, 48, 24), (12, 12), 700


def _rho(seed=3):
    return np.random.default_rng(seed).uniform(0, 1, DESIGN)


@pytest.fixture(scope="module")
def bend(tmp_path_factory, pkg):
    path = tmp_path_factory.mktemp("design") / "bend_48x48x24"
    pkg.cases.bend_case(path, _rho(), shape=SHAPE, spacing=40e-9, steps=STEPS, pml=6,
                        design=DESIGN, pulse_duration=20e-15)
    return path


def _oracle_T(case, rho, box):
    o = OracleSolver(case, dtype=np.float64)
    o.attach()
    o.set_density(box, rho, EPS_LO, EPS_HI)
    o.run()
    return o.dft_transmission(0)


def test_density_forward_writes_the_bundle_bits(pkg, gpu, bend):
    """The HIP map of rho reproduces the bundle written by the host generator."""
    s = pkg.solver.Solver(bend)
    want = s.eps.numpy().copy()
    box = pkg.design._design_box(bend)
    # the box cells start at 1, the others keep the bundle's values
    s.eps.assign(np.where(_box_mask(SHAPE, box).ravel(), 1.0, want).astype(np.float32))
    pkg.design.DensityMap(box).write(_rho(), s)
    np.testing.assert_array_equal(s.eps.numpy(), want)


def _box_mask(shape, box):
    x0, y0, z0, sx, sy, sz = box
    m = np.zeros(shape, bool)
    m[x0:x0 + sx, y0:y0 + sy, z0:z0 + sz] = True
    return m


def test_density_pullback_is_the_derivative_of_the_map(pkg, gpu):
    """pullback(g) = sum_z g * d(inv_eps)/d(rho), checked against central
    differences of the float64 map and against linearity in g."""
    box = (0, 0, 0, 7, 5, 3)
    dm = pkg.design.DensityMap(box)
    rng = np.random.default_rng(0)
    rho = rng.uniform(0, 1, (7, 5))
    g = rng.normal(size=(7, 5, 3))
    got = dm.pullback(rho, torch.from_numpy(g.ravel()).cuda()).cpu().numpy()
    h = 1e-6
    inv = lambda r: 1 / (EPS_LO + r * (EPS_HI - EPS_LO))
    fd = (inv(rho + h) - inv(rho - h)) / (2 * h)
    np.testing.assert_allclose(got, (g * fd[..., None]).sum(-1), rtol=1e-7)
    twice = dm.pullback(rho, torch.from_numpy(2 * g.ravel()).cuda()).cpu().numpy()
    np.testing.assert_array_equal(twice, 2 * got)
    with pytest.raises(ValueError):
        dm.pullback(rho + 1, torch.from_numpy(g.ravel()).cuda())


def test_adjoint_gradient_matches_fp64_finite_differences(pkg, gpu, bend):
    """dT/drho from the checkpointed DFT-plane adjoint against central
    differences (h = 1e-3) of the FP64 oracle on 16 seeded voxels: < 1%."""
    problem = pkg.design.TransmissionProblem(bend, checkpoint_every=64)
    rho = _rho()
    T, grad = problem.value_and_grad(rho)
    grad = grad.cpu().numpy()
    box = problem.map.box
    T64 = _oracle_T(bend, rho, box)
    assert abs(T - T64) / T64 < 1e-5
    rng = np.random.default_rng(16)
    picks = rng.choice(DESIGN[0] * DESIGN[1], 16, replace=False)
    fd, adj = [], []
    for p in picks:
        i, j = divmod(int(p), DESIGN[1])
        up, dn = rho.copy(), rho.copy()
        up[i, j] += 1e-3
        dn[i, j] -= 1e-3
        fd.append((_oracle_T(bend, up, box) - _oracle_T(bend, dn, box)) / 2e-3)
        adj.append(grad[i, j])
    fd, adj = np.array(fd), np.array(adj)
    assert np.linalg.norm(adj - fd) / np.linalg.norm(fd) < 0.01, (adj, fd)
    big = np.abs(fd) > 0.1 * np.abs(fd).max()
    assert np.all(np.abs(adj - fd)[big] < 0.01 * np.abs(fd)[big]), (adj, fd)


def test_autograd_function_and_checkpointing(pkg, gpu, bend):
    """transmission() under torch.autograd gives the problem's own value and
    gradient; the full-history adjoint gives the same bits as checkpoints."""
    ck = pkg.design.TransmissionProblem(bend, checkpoint_every=64)
    full = pkg.design.TransmissionProblem(bend, checkpoint_every=None)
    rho = torch.tensor(_rho(5), dtype=torch.float64, device="cuda", requires_grad=True)
    T = pkg.design.transmission(rho, ck)
    T.backward()
    T_full, g_full = full.value_and_grad(rho.detach())
    assert float(T) == T_full
    np.testing.assert_array_equal(rho.grad.cpu().numpy(), g_full.cpu().numpy())
    # a second design through the same graphs
    rho2 = torch.tensor(_rho(6), dtype=torch.float64, device="cuda", requires_grad=True)
    (2 * pkg.design.transmission(rho2, ck)).backward()
    _, g2 = full.value_and_grad(rho2.detach())
    np.testing.assert_array_equal(rho2.grad.cpu().numpy(), 2 * g2.cpu().numpy())
