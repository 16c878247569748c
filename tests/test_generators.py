"""Synthetic-case generators pinned to the reference bundles (CPU only)."""

import json

import numpy as np
import pytest

from conftest import BUNDLES


@pytest.mark.parametrize("case", ["forward50_gpu", "hann_200fs", "gradient20_highest_exact"])
def test_cpml_rows_match_every_bundle_slab(pkg, case):
    """Roden-Gedney rows with the bundles' cell offsets reproduce every
    exported pml_a/pml_b/inv_kappa array (float32 rounding)."""
    data = np.load(BUNDLES / case / "inputs.npz")
    meta = json.loads((BUNDLES / case / "manifest.json").read_text())
    dt = meta["dt_s"]
    f = next(o for o in meta["objects"] if o["class"] == "ModePlaneSource")["frequency_hz"]
    spacing = meta["settings"]["spacing_nm"] * 1e-9
    checked = 0
    for obj in meta["objects"]:
        if obj["class"] != "PerfectlyMatchedLayer":
            continue
        dims = data[obj["name"] + "/pml_a_E"].shape
        L = max(dims)
        for field in "EH":
            a, b, ik, smax, amax = pkg.cases.cpml_rows(L, spacing, dt, obj["direction"], field, f)
            np.testing.assert_allclose(a, data[f"{obj['name']}/pml_a_{field}"].ravel(), atol=3e-7)
            np.testing.assert_allclose(b, data[f"{obj['name']}/pml_b_{field}"].ravel(), atol=3e-7)
            np.testing.assert_array_equal(ik, data[f"{obj['name']}/inv_kappa_{field}"].ravel())
            assert abs(smax / obj["sigma_end"] - 1) < 1e-6
            assert abs(amax / obj["alpha_start"] - 1) < 1e-6
            checked += 1
    assert checked == 12


def test_mode_solver_matches_bundle_mode(pkg):
    """The eigenmode of the 400x400 nm core at 20 nm: neff and the E/H source
    profiles of source/_E,_H,_neff (FDTDX/tidy3d) up to the sign FDTDX folds
    into profile_phase_shift = pi."""
    data = np.load(BUNDLES / "hann_200fs/inputs.npz")
    meta = json.loads((BUNDLES / "hann_200fs/manifest.json").read_text())
    src = next(o for o in meta["objects"] if o["class"] == "ModePlaneSource")
    eps = 1 / data["inv_permittivity"][0, src["grid_slice"][0][0]].astype(np.float64)
    mode = pkg.modes.solve(eps, 20e-9, 299792458.0 / src["frequency_hz"])[0]
    assert abs(mode.neff / data["source/_neff"].real - 1) < 1e-5
    for field, ours in (("E", mode.E), ("H", mode.H)):
        ref = data[f"source/_{field}"][:, 0]
        for c in (1, 2):
            a = -ours[c].real
            assert np.linalg.norm(a - ref[c]) / np.linalg.norm(ref[c]) < 5e-3


def test_mode_solver_plane_wave_and_power(pkg):
    """A homogeneous section between the zero-padded walls guides its lowest
    mode just below n (walls 2.4 um apart: the continuum TE10 estimate is
    sqrt(n^2 - (lambda/2a)^2) = 1.465); every mode carries unit power
    0.5 sum(Ey Hz - Ez Hy)."""
    eps = np.full((120, 100), 2.25)
    mode = pkg.modes.solve(eps, 20e-9, 1.55e-6)[0]
    assert 1.44 < mode.neff < 1.5
    section = pkg.cases.waveguide_cross_section(96, 80, 20e-9)
    te = pkg.modes.solve(section, 20e-9, 1.55e-6, polarization="TE")[0]
    power = 0.5 * np.sum(te.E[1].real * te.H[2].real - te.E[2].real * te.H[1].real)
    assert abs(power - 1) < 1e-9
    assert 2.35 < te.neff < 2.55
    # a y-normal plane (axis 1) of the same guide rotated gives the same neff
    rotated = pkg.modes.solve_plane(section, 1, 20e-9, 1.55e-6)
    assert rotated.neff > 1.0


def test_bundles_round_trip_through_the_loader(pkg, tmp_path):
    """Synthetic bundles carry a valid sha256 and the reference's layout."""
    path = pkg.cases.slab_case(tmp_path / "slab", n=24, steps=40, pml=4, thickness=6)
    p, meta, data = pkg.solver.load_case(path)
    assert data["inv_permittivity"].shape == (1, 24, 24, 24)
    classes = [o["class"] for o in meta["objects"]]
    assert classes.count("PerfectlyMatchedLayer") == 6
    assert classes.count("DFTPlaneDetector") == 2
    (path / "inputs.npz").write_bytes((path / "inputs.npz").read_bytes() + b"x")
    with pytest.raises(ValueError, match="checksum"):
        pkg.solver.load_case(path)
