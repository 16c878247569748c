"""wfd_resident_run_noalloc (the resident step with the library's own scratch
instead of a caller workspace) against wfd_resident_run, called through the
C-ABI: fields, CPML psi and step counters bit for bit, on two calls in a
row, the second on a smaller grid (the scratch keeps the first call's
contents, which the second must not see).
"""

import ctypes
import sys

import numpy as np
import pytest
import torch

from conftest import BUNDLES

pytestmark = pytest.mark.gpu


def _say(text):
    print(f"[no_alloc] {text}", file=sys.stderr, flush=True)


def _solver(pkg, case):
    s = pkg.solver.Solver(case, resident=False)
    s.compile()
    s.reset()
    return s


def _compare(pkg, case, steps, chunks):
    """The same resident layout stepped `steps` steps in `chunks` launches by
    each entry point; returns the no-allocation run's layout."""
    lib = pkg._lib
    ref = _solver(pkg, case)
    ref.use_resident()
    layout = ref.layout
    got = _solver(pkg, case)
    status = torch.full((1,), -1, dtype=torch.int32, device="cuda:0")
    per = steps // chunks
    for k in range(chunks):
        n = per if k + 1 < chunks else steps - per * (chunks - 1)
        ref.resident_steps(n, record=False)
        plan = got._plan(stage=False)
        lib.call("wfd_resident_run_noalloc", ctypes.byref(plan), ctypes.byref(layout), int(n), 0,
                 ctypes.c_void_p(status.data_ptr()), lib.stream())
    ref.check()
    torch.cuda.synchronize()
    assert int(status.item()) == 0
    for name in ("E", "H", "psi_e", "psi_h", "counters"):
        np.testing.assert_array_equal(getattr(got, name).numpy(), getattr(ref, name).numpy(),
                                      err_msg=f"{case.name} {name}")
    assert int(got.counters.numpy()[0]) == steps
    assert np.abs(got.E.numpy()).max() > 0
    ref.close()
    got.close()
    return layout


def test_noalloc_equals_workspace_run_on_two_calls(pkg, gpu, tmp_path):
    """forward50_gpu (80x80x30, source on) for 300 steps in three launches,
    then a smaller 48x40x32 guide for 150 steps in two: each bitwise equal to
    wfd_resident_run with a zeroed caller workspace."""
    _say("call 1: forward50_gpu 80x80x30, 300 steps")
    big = _compare(pkg, BUNDLES / "forward50_gpu", 300, 3)
    small_case = tmp_path / "guide_48x40x32"
    pkg.cases.waveguide_case(small_case, shape=(48, 40, 32), steps=150, pml=6)
    _say("call 2: guide 48x40x32, 150 steps")
    small = _compare(pkg, small_case, 150, 2)
    assert small.workspace_bytes < big.workspace_bytes


def test_noalloc_rejects_layouts_beyond_the_scratch(pkg, gpu):
    """A layout needing more than the 16 MiB scratch is refused on the host
    (nothing is launched)."""
    lib = pkg._lib
    s = _solver(pkg, BUNDLES / "forward50_gpu")
    layout = s.resident_layout()
    too_big = lib.ResidentLayout.from_buffer_copy(layout)
    too_big.workspace_bytes = (16 << 20) + 8
    status = torch.zeros((1,), dtype=torch.int32, device="cuda:0")
    plan = s._plan(stage=False)
    with pytest.raises(ValueError, match="scratch"):
        lib.call("wfd_resident_run_noalloc", ctypes.byref(plan), ctypes.byref(too_big), 10, 0,
                 ctypes.c_void_p(status.data_ptr()), lib.stream())
    s.close()
