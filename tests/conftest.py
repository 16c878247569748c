import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"
BUNDLES = GOLDEN / "bundles"
EXPECTED = GOLDEN / "expected"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libwfdtd.so")
    config.addinivalue_line("markers", "slow: long CPU oracle run (minutes)")


@pytest.fixture(scope="session")
def pkg():
    import __graft_entry__

    return __graft_entry__.package()


@pytest.fixture(scope="session")
def gpu(pkg):
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("GPU tests need a HIP device; run with -m 'not gpu' on CPU hosts")
    pkg._lib.lib()
    return torch.device("cuda:0")


def rel_l2(a, b):
    import numpy as np

    kind = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a = np.asarray(a, kind)
    b = np.asarray(b, kind)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))
