"""HIP solver path against the CPU oracle (bitwise) and the Warp fixtures."""

import json

import numpy as np
import pytest
import torch

from conftest import BUNDLES, EXPECTED, rel_l2
from oracle import kernels as OK
from oracle.model import OracleSolver

pytestmark = pytest.mark.gpu


def _dev(a, dtype=None):
    return torch.from_numpy(np.ascontiguousarray(a, dtype)).cuda()


def _random_case(seed, shape, packed):
    """Asymmetric, sparse CPML slabs (test_packed.py:12-64) plus a source plane."""
    from oracle.model import pack_coefficients

    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    fields = rng.normal(size=(2, 3, n)).astype("f") * 0.1
    eps = rng.uniform(0.1, 1, n).astype("f")
    coeffs, sizes = [], []
    for field in range(2):
        c = np.zeros((3, max(shape), 3), "f")
        c[..., 1:] = 1
        for axis, length in enumerate(shape):
            active = [0, length - 1] if field == 0 else [1, 2, length - 2]
            if axis == 1:
                active = active[:1]
            c[axis, active] = [-0.12, 0.8, 0.9] if field == 0 else [-0.1, 0.75, 1.1]
        if packed:
            c, size = pack_coefficients(c, shape)
        else:
            size = n
        coeffs.append(c)
        sizes.append(size)
    steps = 24
    source = rng.normal(size=(2, 3, shape[1] * shape[2])).astype("f")
    waves = rng.normal(size=(2, 3, steps)).astype("f")
    return fields, eps, coeffs, sizes, source, waves, steps


@pytest.mark.parametrize("packed", [False, True])
def test_kernel_abi_bitwise_equals_oracle(pkg, gpu, packed):
    """update_e/update_h (+packed) through wp.launch-style calls equal the
    oracle bit for bit over 20 steps, including psi and hprev."""
    k = pkg.kernels
    shape = (7, 8, 9)
    n = int(np.prod(shape))
    fields, eps, coeffs, sizes, source, waves, steps = _random_case(8, shape, packed)
    # oracle state
    oe, oh = fields[0].copy(), fields[1].copy()
    opsi = [np.zeros((6, s), "f") for s in sizes]
    ohp = np.zeros((3, n), "f")
    ostep = np.zeros(1, np.int32)
    # device state
    de, dh = _dev(fields[0]), _dev(fields[1])
    dpsi = [torch.zeros((6, s), device="cuda") for s in sizes]
    dhp = torch.zeros((3, n), device="cuda")
    dstep = torch.zeros(1, dtype=torch.int32, device="cuda")
    deps = _dev(eps)
    dco = [_dev(c) for c in coeffs]
    dsrc = [_dev(s) for s in source]
    dwav = [_dev(w) for w in waves]
    ue = k.update_e_packed if packed else k.update_e
    uh = k.update_h_packed if packed else k.update_h
    save_from, sx, courant = 5, 3, 0.4
    for _ in range(20):
        common = [dhp, save_from, *shape, sx, courant, 0]
        k.launch(ue, n, inputs=[de, dh, deps, dpsi[0], dco[0], dsrc[1], dwav[1], dstep, *common])
        k.launch(uh, n, inputs=[dh, de, deps, dpsi[1], dco[1], dsrc[0], dwav[0], dstep, *common])
        k.launch(k.advance, 1, inputs=[dstep])
        OK.update(True, oe, oh, eps, opsi[0], coeffs[0], source[1], waves[1], ostep, ohp,
                  save_from, shape, sx, courant, 0, packed)
        OK.update(False, oh, oe, eps, opsi[1], coeffs[1], source[0], waves[0], ostep, ohp,
                  save_from, shape, sx, courant, 0, packed)
        ostep[0] += 1
    torch.cuda.synchronize()
    np.testing.assert_array_equal(de.cpu().numpy(), oe)
    np.testing.assert_array_equal(dh.cpu().numpy(), oh)
    np.testing.assert_array_equal(dhp.cpu().numpy(), ohp)
    for d, o in zip(dpsi, opsi):
        np.testing.assert_array_equal(d.cpu().numpy(), o)
    assert int(dstep.item()) == 20


def test_periodic_kernel_abi_bitwise(pkg, gpu):
    """The periodic branch (periodic.py:136-163 reuse) with dummy CPML."""
    k = pkg.kernels
    shape = (4, 4, 16)
    n = int(np.prod(shape))
    rng = np.random.default_rng(3)
    state = rng.normal(size=(6, n)).astype("f")
    coeff = np.zeros((3, 16, 3), "f")
    coeff[:, :, 1:] = 1
    ds = _dev(state)
    deps, dpsi, dco = _dev(np.ones(n, "f")), torch.zeros((1, 1), device="cuda"), _dev(coeff)
    dummy, dstep = torch.zeros((3, 1), device="cuda"), torch.zeros(1, dtype=torch.int32,
                                                                    device="cuda")
    o = state.copy()
    oE, oH = o[:3], o[3:]
    oE, oH = np.ascontiguousarray(oE), np.ascontiguousarray(oH)
    for _ in range(10):
        args = [deps, dpsi, dco, dummy, dummy, dstep, dummy, 1000, *shape, -1, 0.3, 1]
        k.update_e(ds[:3], ds[3:], *args)
        k.update_h(ds[3:], ds[:3], *args)
        oargs = (np.ones(n, "f"), np.zeros((1, 1), "f"), coeff, np.zeros((3, 1), "f"),
                 np.zeros((3, 1), "f"), np.zeros(1, np.int32), np.zeros((3, 1), "f"), 1000,
                 shape, -1, 0.3, 1)
        OK.update(True, oE, oH, *oargs)
        OK.update(False, oH, oE, *oargs)
    np.testing.assert_array_equal(ds[:3].cpu().numpy(), oE)
    np.testing.assert_array_equal(ds[3:].cpu().numpy(), oH)


def test_fused_step_bitwise_equals_oracle(pkg, gpu):
    """The fused column-marching step + all-monitor record equals the oracle's
    restatement of solver.launch_step bit for bit (fields, psi, counters)."""
    s = pkg.solver.Solver(BUNDLES / "forward50_gpu")
    pkg.detectors.attach(s)
    o = OracleSolver(BUNDLES / "forward50_gpu")
    o.attach()
    s.reset()
    for _ in range(150):
        s.launch_step()
        o.launch_step()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(s.E.numpy(), o.E)
    np.testing.assert_array_equal(s.H.numpy(), o.H)
    np.testing.assert_array_equal(s.psi_e.numpy(), o.psi["E"])
    np.testing.assert_array_equal(s.psi_h.numpy(), o.psi["H"])
    assert int(s.step.numpy()[0]) == 150


def test_source_staggering(pkg, gpu):
    """test_warp_solver.py:18-25: after one step E == 0 (ramp 0 at n=0) while
    H, sampled at n+1/2, is nonzero."""
    s = pkg.solver.Solver(BUNDLES / "forward50_gpu", graph_steps=64)
    pkg.detectors.attach(s)
    s.reset()
    s.launch_step()
    torch.cuda.synchronize()
    assert float(np.max(np.abs(s.E.numpy()))) == 0
    assert float(np.max(np.abs(s.H.numpy()))) > 1e-5


def test_full_run_warp_fixture_flux_and_replay(pkg, gpu):
    """test_warp_solver.py:27-48: full 3-D fields and flux traces of
    forward50_gpu, bitwise replay after reset, step counter == steps."""
    s = pkg.solver.Solver(BUNDLES / "forward50_gpu", graph_steps=64)
    pkg.detectors.attach(s)
    s.compile()
    s.reset()
    s.run()
    a = s.fields()
    assert int(s.step.numpy()[0]) == s.steps
    warp = np.load(EXPECTED / "forward50_gpu/warp_fields.npz")
    fdtdx = np.load(EXPECTED / "forward50_gpu/fdtdx_flux.npz")
    for name in ("E", "H"):
        assert np.isfinite(a[name]).all()
        assert rel_l2(a[name], warp[name]) < 3e-6
    for m in s.monitors:
        assert rel_l2(m.flux.numpy(), fdtdx[m.name].ravel()) < 1e-5
    s.reset()
    s.run()
    b = s.fields()
    for name in ("E", "H"):
        np.testing.assert_array_equal(a[name], b[name])


def test_packed_storage_equals_full(pkg, gpu):
    """test_packed.py:93-118: packed CPML storage is a storage change only."""
    full = pkg.solver.Solver(BUNDLES / "forward50_gpu")
    packed = pkg.solver.Solver(BUNDLES / "forward50_gpu", cpml_storage="packed")
    for s in (full, packed):
        pkg.detectors.attach(s)
        s.compile()
        s.reset()
        s.run()
    for name in ("E", "H"):
        np.testing.assert_array_equal(full.fields()[name], packed.fields()[name])
    for a, b in zip(full.monitors, packed.monitors):
        np.testing.assert_array_equal(a.flux.numpy(), b.flux.numpy())
    assert packed.psi_e.capacity + packed.psi_h.capacity < 12 * packed.n * 4
    assert pkg.solver.owned_array_bytes(packed) > packed.E.capacity


def test_staged_replay_equals_per_step_recording(pkg, gpu):
    """hann_200fs: the fused path (footprints staged in the step kernel,
    detectors replayed once per 64-step chunk) and the packed path (every
    monitor recorded each step) give bitwise equal phasors and fluxes."""
    runs = []
    for storage in ("full", "packed"):
        s = pkg.solver.Solver(BUNDLES / "hann_200fs", cpml_storage=storage)
        pkg.detectors.attach(s)
        s.compile()
        s.reset()
        s.run()
        assert len(s.staged()) == (5 if storage == "full" else 0)
        runs.append(s)
    for a, b in zip(runs[0].monitors, runs[1].monitors):
        np.testing.assert_array_equal(a.phasor.numpy(), b.phasor.numpy())
        np.testing.assert_array_equal(a.flux.numpy(), b.flux.numpy())
    np.testing.assert_array_equal(runs[0].E.numpy(), runs[1].E.numpy())


def test_hann20_modal_phasors_and_transmission(pkg, gpu):
    """3M-cell bend, 5,245 steps, exact colocation + Hann-windowed modal
    phasors: against the Warp phasors and transmission (Warp vs FDTDX: T abs
    2.4e-7, E/H 3.6e-6)."""
    s = pkg.solver.Solver(BUNDLES / "hann_200fs")
    pkg.detectors.attach(s)
    s.compile()
    s.reset()
    s.run()
    got = pkg.detectors.measurements(s)
    ref = json.loads((EXPECTED / "hann20/measurements.json").read_text())["warp"]
    assert abs(got["modal_transmission"] - ref["modal_transmission"]) < 2e-6
    for m in s.monitors:
        if m.is_phasor:
            want = np.load(EXPECTED / f"hann20/{m.name}_phasors.npz")["phasors"]
            assert rel_l2(m.phasors(), want) < 1e-5
    f = s.fields()
    warp = np.load(EXPECTED / "hann20/warp_fields_midz.npz")
    for name in ("E", "H"):
        mid = f[name][..., f[name].shape[-1] // 2]
        assert rel_l2(mid, warp[name]) < 1e-5
