"""bench.py's contract pieces that need no GPU: the roofline's algorithmic
bytes per cell-update and the refusal of more than one GPU."""

import subprocess
import sys

import pytest

from conftest import ROOT


@pytest.fixture(scope="module")
def bench():
    sys.path.insert(0, str(ROOT))
    import bench as module

    return module


@pytest.mark.parametrize("shape, pml, f_cpml, B", [
    ((160, 96, 80), 12, 0.70, 74.4),    # straight waveguide, 12-cell CPML
    ((64, 64, 64), 10, 0.9375, 82.0),   # 64^3 slab, 10-cell CPML
    ((200, 200, 80), 10, 0.45, 66.4),   # 200x200x80 bend, 10-cell CPML
])
def test_algorithmic_bytes_per_cell_update(bench, shape, pml, f_cpml, B):
    """52 B of fields and 1/eps plus 32 B of psi traffic per CPML axis a
    cell lies in, averaged over the grid (f_cpml = sum over axes of 2L/N)."""
    b, f = bench.algorithmic_bytes_per_cell(shape, pml)
    assert f == pytest.approx(f_cpml, abs=1e-12)
    assert b == pytest.approx(B, abs=1e-9)


def test_more_than_one_gpu_is_refused():
    out = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert out.returncode == 2
    assert "one GPU" in out.stderr and out.stdout == ""
