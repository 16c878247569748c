"""The modal overlap of the DFT planes separates propagation directions.

mode_amplitudes() projects a plane's phasors on the stored forward mode with
alpha = sum((E_m x H* + E* x H_m)[axis]) / 4.  A field equal to the forward
mode must give |alpha|^2 = (Re sum (E_m x H_m*)[axis] / 2)^2 > 0, and the same
mode travelling backwards (transverse E unchanged, transverse H negated,
longitudinal components the other way round) must project to ~0 -- a modal
transmission that counts backward power would not be bounded by 1 on a
passive bend.  Checked on the bend bundle's own mode_in / mode_out fields at
every wavelength (CPU; phasors substituted for a solver run).
"""

import types

import numpy as np
import pytest

N = 9  # wavelengths of the surrogate bundle: 8 spectrum + the objective's


@pytest.fixture(scope="module")
def bend(tmp_path_factory, pkg):
    path = tmp_path_factory.mktemp("projection") / "bend"
    pkg.surrogate.surrogate_case(path, side=24, shape=(64, 64, 40), spacing=40e-9, steps=40,
                                 pml=6)
    _, meta, data = pkg.solver.load_case(path)
    return meta, data


def _monitor(pkg, meta, data, name):
    obj = next(o for o in meta["objects"] if o["name"] == name)
    stub = types.SimpleNamespace(steps=int(meta["time_steps"]), meta=meta, data=data,
                                 device="cpu", shape=data["inv_permittivity"].shape[-3:])
    return pkg.detectors.DFTMonitor(stub, obj)


@pytest.mark.parametrize("name", ["mode_in", "mode_out"])
def test_backward_mode_projects_to_zero(pkg, bend, name):
    meta, data = bend
    m = _monitor(pkg, meta, data, name)
    assert m.nf == N
    em = np.asarray(data[name + "/_mode_E"], np.complex128)  # (nf, 3, *plane)
    hm = np.asarray(data[name + "/_mode_H"], np.complex128)
    a = m.axis
    flip = np.ones((1, 3) + (1,) * (em.ndim - 2))
    flip[0, a] = -1  # longitudinal component reverses with the direction
    forward = np.concatenate([em, hm], axis=1)
    backward = np.concatenate([em * flip, -hm * flip], axis=1)
    m.phasors = lambda: forward
    p_fwd = m.power()
    want = (np.real(np.sum(np.cross(em, hm.conj(), axis=1)[:, a],
                           axis=tuple(range(1, em.ndim - 1)))) / 2) ** 2
    np.testing.assert_allclose(p_fwd, want, rtol=1e-10)
    assert np.all(p_fwd > 0)
    m.phasors = lambda: backward
    p_bwd = m.power()
    assert np.all(p_bwd <= 1e-8 * p_fwd), p_bwd / p_fwd
    # a planted sign error (H not negated) projects fully: the check has teeth
    m.phasors = lambda: np.concatenate([em * flip, hm * flip], axis=1)
    assert np.all(m.power() > 0.5 * p_fwd)
