"""The CPU oracle against the reference's own Warp fixtures (forward path)."""

import numpy as np
import pytest

from conftest import BUNDLES, EXPECTED, rel_l2
from oracle.model import OracleSolver


@pytest.fixture(scope="module")
def forward50():
    runs = {}
    for dtype in (np.float32, np.float64):
        s = OracleSolver(BUNDLES / "forward50_gpu", dtype=dtype)
        s.attach()
        s.run()
        runs[dtype] = s
    return runs


def test_fields_match_warp_full_3d(forward50):
    """test_warp_solver.py:27-41 at the fixture's own precision: the reference
    Warp fields after 2,098 steps (Warp vs FDTDX is 5.0e-6 / 4.3e-6)."""
    warp = np.load(EXPECTED / "forward50_gpu/warp_fields.npz")
    s = forward50[np.float32]
    f = s.fields()
    assert int(s.step[0]) == s.steps
    for name in ("E", "H"):
        assert np.isfinite(f[name]).all()
        assert rel_l2(f[name], warp[name]) < 3e-6


def test_flux_traces_match_fdtdx(forward50):
    ref = np.load(EXPECTED / "forward50_gpu/fdtdx_flux.npz")
    for m in forward50[np.float32].monitors:
        assert rel_l2(m.flux, ref[m.name].ravel()) < 1e-5


def test_float32_vs_float64_trajectory(forward50):
    """Rounding error of the float32 restatement against its float64 twin; the
    Warp fixture sits at the same distance from the float64 run."""
    f32 = forward50[np.float32].fields()
    f64 = forward50[np.float64].fields()
    warp = np.load(EXPECTED / "forward50_gpu/warp_fields.npz")
    for name in ("E", "H"):
        assert rel_l2(f32[name], f64[name]) < 2e-6
        assert rel_l2(warp[name], f64[name]) < 2e-6


def test_planted_sign_error_is_caught():
    """A flipped source sign in the oracle must fail the fixture check."""
    s = OracleSolver(BUNDLES / "forward50_gpu")
    s.sources["H"] = -s.sources["H"]
    s.run(300)
    o = OracleSolver(BUNDLES / "forward50_gpu")
    o.run(300)
    assert rel_l2(s.fields()["E"], o.fields()["E"]) > 0.1
