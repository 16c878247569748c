"""Rounding error of the float32 path at the headline workload, against the
float64 twin of the CPU oracle (the GPU path equals the float32 oracle bit
for bit: test_gpu_atscale.py).

bench.py's 160x96x80 waveguide (12-cell CPML, TE0 Hann source, 8-wavelength
DFT planes) over all 4000 steps in float32 and in float64: the fields and
the FP64-accumulated DFT phasors of both planes (from which the modal
transmission follows) must agree within the 1e-5 relative L2 of the float32
contract.  About four minutes of CPU time, so it runs only with
WFDTD_SLOW=1.
"""

import os
import sys

import numpy as np
import pytest

from conftest import ROOT, rel_l2
from oracle.model import OracleSolver

pytestmark = [pytest.mark.slow,
              pytest.mark.skipif(not os.environ.get("WFDTD_SLOW"),
                                 reason="four minutes of CPU time: set WFDTD_SLOW=1")]


def test_waveguide_float32_vs_float64():
    """While the pulse is in the guide (step 1600) the fields agree within
    1e-6; after 4000 steps the pulse has left through the CPML (the field
    norm falls about 2600-fold) and what remains is compared against the
    largest field norm of the run; the DFT phasors, which integrate the
    whole pulse, agree within 1e-6."""
    sys.path.insert(0, str(ROOT))
    import bench

    case = bench.build_case()[1]
    runs = [OracleSolver(case, dtype=dtype) for dtype in (np.float32, np.float64)]
    for s in runs:
        s.attach()
    a, b = runs
    errors, peak = {}, {"E": 0.0, "H": 0.0}
    done = 0
    for upto in (800, 1600, 2400, 4000):
        for s in runs:
            s.run(upto - done)
        done = upto
        for name in ("E", "H"):
            peak[name] = max(peak[name], float(np.linalg.norm(getattr(b, name))))
            if upto == 1600:
                errors[f"{name}@1600"] = rel_l2(getattr(a, name), getattr(b, name))
    for name in ("E", "H"):
        diff = np.linalg.norm(getattr(a, name).astype(np.float64) - getattr(b, name))
        errors[f"{name}@4000/peak"] = float(diff / peak[name])
    for ma, mb in zip(a.monitors, b.monitors):
        errors[ma.name] = rel_l2(ma.phasor, mb.phasor)
    print(f"float32 vs float64: {errors}", file=sys.stderr)
    assert max(errors.values()) < 1e-5, errors
