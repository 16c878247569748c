"""The benchmarked and batched workloads at their own shapes, and the
reference's long-horizon fixtures.

* The headline workload exactly as bench.py runs it (160x96x80 waveguide,
  12-cell CPML, TE0 Hann source, 8-wavelength DFT planes, resident kernel in
  1000-step launches) against the CPU oracle over all 4000 steps.
* The 32-design batch (seeded 120x120x64 .. 256x224x96 bends) against solo
  runs of every design and, for the smallest and largest, the oracle.
* The reference's 24M-cell runs (forward10_gpu: 10,490 steps; hann_10nm:
  15,735 steps with Hann-windowed modal phasors) and the 3M-cell bend with
  the flow-plane PhasorDetector (bend_flow20: 7,868 steps) against the Warp
  outputs the reference committed (results/warp/{forward10_gpu, hann10,
  bend_flow20}); Warp against FDTDX differs by 6.3e-6 / 3.1e-6 (forward10),
  7.1e-6 / 5.7e-6 (hann10) and 6.3e-6 / 5.3e-6 (bend_flow20) in E / H.
"""

import json
import sys
import tempfile
from pathlib import Path

import numpy as np
import pytest
import torch

from conftest import BUNDLES, EXPECTED, ROOT, rel_l2
from oracle.model import OracleSolver

pytestmark = pytest.mark.gpu


def _say(text):
    print(f"[atscale] {text}", file=sys.stderr, flush=True)


def _bench_waveguide(pkg):
    """The bundle bench.py writes and times (same path, same generator)."""
    sys.path.insert(0, str(ROOT))
    import bench

    return bench.build_case()[1]


def test_benchmark_waveguide_resident_bitwise_vs_oracle(pkg, gpu):
    """bench.py's case on the resident kernel with its 5x12x80 sub-boxes and
    1000-step launches: E, H, psi and counters bitwise equal to the oracle
    after 4000 steps; the FP64 DFT planes within 1e-12 (the replay sums the
    steps of a chunk in eight interleaved partial sums, the oracle in one)."""
    case = _bench_waveguide(pkg)
    s = pkg.solver.Solver(case, resident=True, graph_steps=1000)
    pkg.detectors.attach(s)
    s.compile()
    assert s.resident
    L = s.layout
    assert (L.sx, L.sy, L.sz, L.blocks) == (5, 12, 80, 256)
    s.reset()
    launches = []
    s.run(events=launches)
    assert [n for _, _, n in launches] == [1000] * 4
    _say("GPU run done; oracle 4000 steps")
    o = OracleSolver(case)
    o.attach()
    o.run()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(s.E.numpy(), o.E)
    np.testing.assert_array_equal(s.H.numpy(), o.H)
    np.testing.assert_array_equal(s.psi_e.numpy(), o.psi["E"])
    np.testing.assert_array_equal(s.psi_h.numpy(), o.psi["H"])
    assert int(s.step.numpy()[0]) == 4000
    assert len(s.monitors) == len(o.monitors) == 2
    for m, om in zip(s.monitors, o.monitors):
        assert m.name == om.name
        assert rel_l2(m.phasor.numpy(), om.phasor) < 1e-12
    T = pkg.detectors.dft_transmission(s)
    assert np.all(np.abs(T - 1) < 0.01)


@pytest.fixture(scope="module")
def batch32(pkg):
    """The seeded 32-design batch bench.py times (bundles shared with it)."""
    root = Path(tempfile.gettempdir()) / "wfdtd_bench_batch32"
    _say("writing / reusing the 32 design bundles")
    return pkg.cases.batch_cases(root, count=32, seed=0, steps=3000)


BATCH_STEPS = 300


def test_batch32_every_design_equals_its_solo_run(pkg, gpu, batch32):
    """All 32 designs (87M cells) stepped together for 300 steps: fields,
    psi, counters and FP64 DFT planes of each equal the same design stepped
    alone, bit for bit; the smallest and largest designs also equal the CPU
    oracle bitwise (fields, psi) and within 1e-12 (DFT planes)."""
    paths = [p for p, _ in batch32]
    solvers = []
    for p in paths:
        s = pkg.solver.Solver(p, resident=False)
        pkg.detectors.attach(s)
        solvers.append(s)
    batch = pkg.batch.Batch(solvers)
    batch.compile()
    batch.reset()
    batch.run_steps(BATCH_STEPS)
    cells = [s.n for s in solvers]
    assert sum(cells) == 87152128
    _say(f"batch of {len(solvers)} stepped {BATCH_STEPS} steps; solo runs")
    for k, (p, b) in enumerate(zip(paths, solvers)):
        solo = pkg.solver.Solver(p, resident=False, x_chunk=b.x_chunk)
        pkg.detectors.attach(solo)
        solo.compile()
        solo.reset()
        solo.run_steps(BATCH_STEPS)
        for name in ("E", "H", "psi_e", "psi_h", "counters"):
            np.testing.assert_array_equal(getattr(solo, name).numpy(), getattr(b, name).numpy(),
                                          err_msg=f"design {k} {name}")
        for ma, mb in zip(solo.monitors, b.monitors):
            np.testing.assert_array_equal(ma.phasor.numpy(), mb.phasor.numpy(),
                                          err_msg=f"design {k} {ma.name}")
        solo.close()
        del solo
    for k in (int(np.argmin(cells)), int(np.argmax(cells))):
        _say(f"oracle {BATCH_STEPS} steps of design {k} ({solvers[k].shape})")
        o = OracleSolver(paths[k])
        o.attach()
        o.run(BATCH_STEPS)
        b = solvers[k]
        np.testing.assert_array_equal(b.E.numpy(), o.E)
        np.testing.assert_array_equal(b.H.numpy(), o.H)
        np.testing.assert_array_equal(b.psi_e.numpy(), o.psi["E"])
        np.testing.assert_array_equal(b.psi_h.numpy(), o.psi["H"])
        for m, om in zip(b.monitors, o.monitors):
            assert m.name == om.name
            assert rel_l2(m.phasor.numpy(), om.phasor) < 1e-12
    batch.close()


def _midz(fields):
    return {k: v[..., v.shape[-1] // 2] for k, v in fields.items()}


def test_forward10_24m_cells_long_horizon(pkg, gpu):
    """400x400x150 at 10 nm, 10,490 steps, two flux planes (the streaming
    kernel pair: the grid is past the resident capacity): mid-z E and H
    within 1e-5 of Warp."""
    s = pkg.solver.Solver(BUNDLES / "forward10_gpu")
    pkg.detectors.attach(s)
    s.compile()
    assert not s.resident and s.n == 24_000_000
    s.reset()
    s.run()
    assert int(s.step.numpy()[0]) == 10490
    got = _midz(s.fields())
    warp = np.load(EXPECTED / "forward10_gpu/warp_fields_midz.npz")
    for name in ("E", "H"):
        assert np.isfinite(got[name]).all()
        err = rel_l2(got[name], warp[name])
        _say(f"forward10 {name} rel-L2 vs Warp {err:.3e}")
        assert err < 1e-5
    s.close()


def test_hann10_24m_cells_modal_transmission(pkg, gpu):
    """400x400x150 bend at 10 nm, 15,735 steps, exact-colocation Hann modal
    phasors: T within 2e-6 of Warp's 0.9319190425 (Warp vs FDTDX: 4.3e-9),
    mid-z E and H within 1e-5."""
    s = pkg.solver.Solver(BUNDLES / "hann_10nm")
    pkg.detectors.attach(s)
    s.compile()
    s.reset()
    s.run()
    got = pkg.detectors.measurements(s)
    ref = json.loads((EXPECTED / "hann10/measurements.json").read_text())["warp"]
    _say(f"hann10 T {got['modal_transmission']:.10f} (Warp {ref['modal_transmission']:.10f})")
    assert abs(got["modal_transmission"] - ref["modal_transmission"]) < 2e-6
    assert abs(got["modal_reflection"] - ref["modal_reflection"]) < 2e-6
    f = _midz(s.fields())
    warp = np.load(EXPECTED / "hann10/warp_fields_midz.npz")
    for name in ("E", "H"):
        err = rel_l2(f[name], warp[name])
        _say(f"hann10 {name} rel-L2 vs Warp {err:.3e}")
        assert err < 1e-5
    s.close()


def test_bend_flow20_flow_plane_phasors(pkg, gpu):
    """The 3M-cell bend with a 200x200 flow-plane PhasorDetector at z = 37
    (staged like the modal planes): flow phasors within 1e-5 of Warp's
    (6, 200, 200, 1) array, T within 2e-6, mid-z fields within 1e-5."""
    s = pkg.solver.Solver(BUNDLES / "bend_flow20")
    pkg.detectors.attach(s)
    s.compile()
    names = [m.name for m in s.staged()]
    assert "flow" in names and len(names) == len(s.monitors)
    s.reset()
    s.run()
    flow = next(m for m in s.monitors if m.name == "flow")
    want = np.load(EXPECTED / "bend_flow20/flow_phasors.npz")["phasors"]
    err = rel_l2(flow.phasors(), want)
    _say(f"flow plane rel-L2 vs Warp {err:.3e}")
    assert flow.phasors().shape == want.shape and err < 1e-5
    got = pkg.detectors.measurements(s)
    ref = json.loads((EXPECTED / "bend_flow20/measurements.json").read_text())["warp"]
    assert abs(got["modal_transmission"] - ref["modal_transmission"]) < 2e-6
    f = _midz(s.fields())
    warp = np.load(EXPECTED / "bend_flow20/warp_fields_midz.npz")
    for name in ("E", "H"):
        assert rel_l2(f[name], warp[name]) < 1e-5
    s.close()
