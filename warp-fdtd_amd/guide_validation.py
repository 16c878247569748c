"""Full-field finiteness checks without copying large grids to the host
(drop-in for warp_fdtd/guide_validation.py)."""

import numpy as np

from ._lib import call, ptr, stream
from .array import Array
from .kernels import Kernel


def _count_nonfinite(field, count):
    n = getattr(field, "tensor", field).numel()
    call("wfd_count_nonfinite", ptr(field), int(n), ptr(count), stream())


# guide_validation.py:6-9 — count += number of non-finite values of field.
count_nonfinite = Kernel("wfd_count_nonfinite", _count_nonfinite)


def assert_finite(solver):
    """guide_validation.py:12-20"""
    count = Array.zeros(1, np.int32, solver.device)
    for field in (solver.e, solver.h):
        _count_nonfinite(field, count)
    n = int(count.numpy()[0])
    if n:
        raise FloatingPointError(f"{n} nonfinite field values")
