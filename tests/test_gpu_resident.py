"""On-chip-resident stepping (wfd_resident_run) against the fused kernel pair.

The resident kernel performs the fused step's arithmetic operation for
operation, so fields, CPML state, step counters and every staged detector
must agree bit for bit with the streaming path (itself bitwise equal to the
CPU oracle, test_gpu_solver.py) over whole runs, ragged chunks and quiet
stretches.
"""

import numpy as np
import pytest
import torch

from conftest import BUNDLES

pytestmark = pytest.mark.gpu

RESIDENT_CELLS = 40896 // 7  # LDS floats per workgroup (kLds) / 7 floats per cell


def _pair(pkg, case, attach=True, **kw):
    out = []
    for resident in (False, True):
        s = pkg.solver.Solver(case, resident=resident, **kw)
        if attach:
            pkg.detectors.attach(s)
        s.compile()
        s.reset()
        out.append(s)
    assert not out[0].resident and out[1].resident
    return out


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
def waveguide(tmp_path_factory, pkg):
    """A small straight guide with modal + flux DFT planes (kind-2 replay)."""
    path = tmp_path_factory.mktemp("resident") / "waveguide_96x48x40"
    pkg.cases.waveguide_case(path, shape=(96, 48, 40), steps=300, pml=8)
    return path


def test_layout_covers_the_grid(pkg, gpu):
    s = pkg.solver.Solver(BUNDLES / "forward50_gpu", resident=False)
    L = s.resident_layout()
    nx, ny, nz = s.shape
    for size, groups, n in ((L.sx, L.gx, nx), (L.sy, L.gy, ny), (L.sz, L.gz, nz)):
        assert size * groups >= n and size * (groups - 1) < n
    assert L.blocks == L.gx * L.gy * L.gz <= L.capacity
    assert L.sx * L.sy * L.sz <= RESIDENT_CELLS
    assert L.workspace_bytes % 4 == 0
    # an explicit sub-box is honoured (evened along each axis)
    M = s.resident_layout(20, 20, 10)
    assert (M.gx, M.gy, M.gz) == (4, 4, 3) and M.sx * 4 >= nx


def test_forward50_run_bitwise(pkg, gpu):
    """Whole 2,098-step run with the two flux planes staged."""
    a, b = _pair(pkg, BUNDLES / "forward50_gpu")
    assert len(b.staged()) == len(b.monitors) == 2
    a.run()
    b.run()
    _assert_same(a, b)
    assert int(b.step.numpy()[0]) == b.steps


def test_dft_waveguide_bitwise(pkg, gpu, waveguide):
    """FP64 multi-wavelength DFT planes replayed from resident staging."""
    a, b = _pair(pkg, waveguide)
    assert all(m.obj["class"] == "DFTPlaneDetector" for m in b.monitors)
    a.run()
    b.run()
    _assert_same(a, b)


def test_ragged_run_steps_bitwise(pkg, gpu, waveguide):
    """run_steps of ragged lengths (chunk tails, ring wrap) agree after each."""
    a, b = _pair(pkg, waveguide)
    for count in (37, 100, 1, 64, 70):
        a.run_steps(count)
        b.run_steps(count)
        _assert_same(a, b)


def test_quiet_single_launch_bitwise(pkg, gpu):
    """Without monitors a run is one launch of every step."""
    a, b = _pair(pkg, BUNDLES / "forward50_gpu", attach=False)
    a.run_steps(777)
    b.run_steps(777)
    _assert_same(a, b)


def test_explicit_layouts_bitwise(pkg, gpu):
    """Other decompositions (ragged edge workgroups) give the same bits."""
    ref = pkg.solver.Solver(BUNDLES / "forward50_gpu", resident=False)
    ref.compile()
    ref.reset()
    ref.run_steps(150)
    for box in ((80, 2, 30), (7, 13, 30), (11, 80, 4)):
        s = pkg.solver.Solver(BUNDLES / "forward50_gpu", resident=False)
        s.compile()
        s.reset()
        s.use_resident(s.resident_layout(*box))
        s.run_steps(150)
        _assert_same(ref, s)


def test_not_applicable_configurations(pkg, gpu):
    """resident=True is refused where the kernel cannot run the step."""
    with pytest.raises(ValueError):  # 3M cells: beyond the co-resident LDS
        s = pkg.solver.Solver(BUNDLES / "hann_200fs", resident=True)
        s.compile()
    s = pkg.solver.Solver(BUNDLES / "forward50_gpu", cpml_storage="packed", resident=True)
    with pytest.raises(ValueError):
        s.compile()
    s = pkg.solver.Solver(BUNDLES / "forward50_gpu", full_hprev=True, resident=True)
    with pytest.raises(ValueError):
        s.compile()
    auto = pkg.solver.Solver(BUNDLES / "hann_200fs")
    pkg.detectors.attach(auto)
    auto.tune()
    assert not auto.resident


def test_bounded_wait_reports_timeout(pkg, gpu):
    """A spin limit of one poll makes neighbour waits give up: every
    workgroup still exits, the error word is raised, check() reports it and
    clears the workspace, and the next run is correct again."""
    a, b = _pair(pkg, BUNDLES / "forward50_gpu", attach=False)
    b.resident_steps(64, record=False, spin_limit=1)
    with pytest.raises(RuntimeError, match="timed out"):
        b.check()
    b.reset()
    a.run_steps(64)
    b.run_steps(64)
    _assert_same(a, b)
