"""Multi-wavelength DFT planes, pulsed sources and synthetic workloads on the GPU."""

import numpy as np
import pytest
import torch

from conftest import rel_l2
from oracle import kernels as OK

pytestmark = pytest.mark.gpu


def _oracle_replay(s, steps):
    """Step the oracle with the solver's own tables and record every DFT plane
    with the oracle's sequential FP64 running DFT."""
    n = s.n
    E = np.zeros((3, n), np.float32)
    H = np.zeros((3, n), np.float32)
    hp = np.zeros((3, n), np.float32)
    eps = s.eps.numpy()
    psi = [np.zeros((6, n), np.float32) for _ in range(2)]
    coeff = [np.ascontiguousarray(s._coeff_rows[f]) for f in "EH"]
    src = {f: s.sources[f].numpy() for f in "EH"}
    wav = {f: s.waves[f].numpy() for f in "EH"}
    step = np.zeros(1, np.int32)
    ph = {m.name: np.zeros((m.nf, 6, m.n, 2)) for m in s.monitors}
    w = {m.name: np.ascontiguousarray(m.weights.numpy()) for m in s.monitors}
    for t in range(steps):
        OK.update(True, E, H, eps, psi[0], coeff[0], src["H"], wav["H"], step, hp,
                  s.save_h_from, s.shape, s.source_x, s.courant, 0)
        OK.update(False, H, E, eps, psi[1], coeff[1], src["E"], wav["E"], step, hp,
                  s.save_h_from, s.shape, s.source_x, s.courant, 0)
        for m in s.monitors:
            OK.record_dft(E, H, hp, t, w[m.name], ph[m.name], s.shape, m.lo, m.shape, m.exact)
        step[0] += 1
    return E, H, ph


def test_dft_planes_match_oracle_across_chunks(pkg, gpu, tmp_path):
    """Graph-chunked stage/accumulate (64-step chunks + a 22-step tail) equals
    the oracle's per-step FP64 running DFT; fields stay bitwise equal."""
    case = pkg.cases.slab_case(tmp_path / "slab", n=24, spacing=50e-9, steps=150, pml=4,
                               thickness=6)
    s = pkg.solver.Solver(case, graph_steps=64)
    pkg.detectors.attach(s)
    s.compile()
    s.reset()
    s.run()
    E, H, ph = _oracle_replay(s, s.steps)
    np.testing.assert_array_equal(s.E.numpy(), E)
    np.testing.assert_array_equal(s.H.numpy(), H)
    for m in s.monitors:
        got = m.phasor.numpy()
        assert rel_l2(got, ph[m.name]) < 1e-12


def test_vacuum_plane_wave_flux_is_conserved(pkg, gpu, tmp_path):
    """Config-1 geometry without the slab: the DFT flux after equals the flux
    before at every wavelength (lossless propagation of the plane pulse)."""
    case = pkg.cases.slab_case(tmp_path / "vac", vacuum=True)
    s = pkg.solver.Solver(case)
    pkg.detectors.attach(s)
    s.compile()
    s.reset()
    s.run()
    m = {x.name: x for x in s.monitors}
    ratio = m["flux_after"].flux() / m["flux_before"].flux()
    assert np.all(np.abs(ratio - 1) < 0.03), ratio


def test_straight_waveguide_modal_transmission(pkg, gpu, tmp_path):
    """Config 2 (160x96x80 at 20 nm, 4000 steps): the TE0 modal source of the
    eigenmode solver propagates with unit modal transmission at all 8
    wavelengths, and the flux DFTs agree with the modal powers."""
    case = pkg.cases.waveguide_case(tmp_path / "wg")
    s = pkg.solver.Solver(case)
    pkg.detectors.attach(s)
    s.compile()
    s.reset()
    s.run()
    m = {x.name: x for x in s.monitors}
    T = m["mode_out"].power() / m["mode_in"].power()
    assert np.all(np.abs(T - 1) < 0.01), T
    flux_ratio = m["mode_out"].flux() / m["mode_in"].flux()
    assert np.all(np.abs(flux_ratio - 1) < 0.02), flux_ratio
    assert np.isfinite(s.E.numpy()).all()


@pytest.mark.parametrize("resident", [False, True])
def test_config1_slab_bitwise_and_dft_against_oracle(pkg, gpu, tmp_path, resident):
    """Config 1 at its full size (64^3, 10-cell CPML, 16-cell slab, 600 steps,
    5 wavelengths): fields bitwise equal to the oracle on both the fused and
    the on-chip-resident path, the FP64 DFT planes within 1e-12."""
    case = pkg.cases.slab_case(tmp_path / "slab")
    s = pkg.solver.Solver(case, resident=resident)
    pkg.detectors.attach(s)
    s.compile()
    assert bool(s.resident) == resident
    s.reset()
    s.run()
    E, H, ph = _oracle_replay(s, s.steps)
    np.testing.assert_array_equal(s.E.numpy(), E)
    np.testing.assert_array_equal(s.H.numpy(), H)
    for m in s.monitors:
        assert rel_l2(m.phasor.numpy(), ph[m.name]) < 1e-12, m.name


def _fabry_perot(wavelengths, n_slab, cells, spacing, courant):
    """Power transmission of a lossless slab at normal incidence with the
    phase thickness of the Yee scheme's 1D numerical wavenumber,
    sin(k dx / 2) = (n / S) sin(w dt / 2), and the continuum Fresnel
    reflectance R = ((n - 1) / (n + 1))^2."""
    lam = np.asarray(wavelengths)
    half_wdt = np.pi * courant * spacing / lam
    k_dx = 2 * np.arcsin(n_slab / courant * np.sin(half_wdt))
    delta = 2 * k_dx * cells
    R = ((n_slab - 1) / (n_slab + 1)) ** 2
    F = 4 * R / (1 - R) ** 2
    return 1 / (1 + F * np.sin(delta / 2) ** 2)


def test_config1_slab_transmission(pkg, gpu, tmp_path):
    """Config-1 geometry run to 3000 steps (the 600-step window ends while the
    slab still rings): transmission = flux after the slab / flux after in
    vacuum; the net flux before the slab over the vacuum run's is 1 - R, which
    for a lossless slab equals T (energy conservation); T follows the
    Fabry-Perot formula of a 16-cell, eps = 12.08 slab."""
    flux = {}
    for name, vacuum in (("slab", False), ("vac", True)):
        s = pkg.solver.Solver(pkg.cases.slab_case(tmp_path / name, vacuum=vacuum, steps=3000))
        pkg.detectors.attach(s)
        s.compile()
        s.reset()
        s.run()
        m = {x.name: x for x in s.monitors}
        flux[name] = (m["flux_before"].flux(), m["flux_after"].flux())
        wl = np.array(m["flux_after"].wavelengths)
    T = flux["slab"][1] / flux["vac"][1]
    one_minus_R = flux["slab"][0] / flux["vac"][0]
    fp = _fabry_perot(wl, np.sqrt(12.08), 16, 50e-9, 0.99 / np.sqrt(3))
    report = f"T={T}, 1-R={one_minus_R}, Fabry-Perot={fp}"
    # measured on MI355X: |T - (1 - R)| <= 3e-4; T - FP from -0.016 to -0.038 (the
    # grid's interfaces and impedances against the continuum Fresnel terms)
    assert np.all(np.abs(T - one_minus_R) < 0.005), report
    assert np.all(np.abs(T - fp) < 0.05), report
