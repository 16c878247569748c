"""ctypes binding of libwfdtd.so (the HIP C-ABI declared in include/wfdtd.h).

The library is loaded from this package directory; there is no fallback: a
missing or unloadable library raises ImportError/OSError at first use, so the
product path never runs anywhere but on the HIP kernels.
"""

import atexit
import ctypes
import itertools
import sys
import weakref
from pathlib import Path

import torch

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libwfdtd.so"

WFD_OK = 0
WFD_ERR_ARGUMENT = -1
WFD_ERR_HIP = -2
WFD_ERR_UNSUPPORTED = -3
WFD_MAX_MONITORS = 16

P = ctypes.c_void_p
I = ctypes.c_int
L = ctypes.c_int64
F = ctypes.c_float


WFD_MAX_STAGES = 8
WFD_MAX_BATCH = 64


class StageDesc(ctypes.Structure):
    """wfd_stage (include/wfdtd.h)."""

    _fields_ = [("buf", P), ("box", I * 6)]


class YeePlan(ctypes.Structure):
    """wfd_yee_plan (include/wfdtd.h)."""

    _fields_ = [
        ("e", P), ("h", P), ("inv_eps", P), ("psi_e", P), ("psi_h", P),
        ("coeff_e", P), ("coeff_h", P), ("src_e", P), ("src_h", P),
        ("wave_e", P), ("wave_h", P), ("counters", P), ("hprev", P),
        ("hbox", I * 6), ("save_h_from", I), ("coeff_rows", I), ("steps", I),
        ("nx", I), ("ny", I), ("nz", I), ("source_x", I), ("courant", F),
        ("x_chunk", I), ("nstage", I), ("slots", I), ("stages", StageDesc * WFD_MAX_STAGES),
    ]


class YeeTargets(ctypes.Structure):
    """wfd_yee_targets (include/wfdtd.h)."""

    _fields_ = [("e", P), ("h", P), ("psi_e", P), ("psi_h", P), ("ticket", P), ("planes", I),
                ("slots_per_thread", I), ("steps", I), ("hist", P), ("hist_steps", I),
                ("hist_box", I * 6), ("threads", I)]


class InjectSource(ctypes.Structure):
    """wfd_inject_source (include/wfdtd.h)."""

    _fields_ = [("source", P), ("weights", P), ("x0", I), ("y0", I), ("z0", I), ("sy", I),
                ("sz", I), ("count", I), ("steps", I)]


WFD_MAX_INJECT = 8


class ResidentLayout(ctypes.Structure):
    """wfd_resident_layout (include/wfdtd.h)."""

    _fields_ = [
        ("sx", I), ("sy", I), ("sz", I), ("gx", I), ("gy", I), ("gz", I),
        ("blocks", I), ("capacity", I), ("workspace_bytes", L),
    ]


class ReplayDesc(ctypes.Structure):
    """wfd_replay (include/wfdtd.h)."""

    _fields_ = [
        ("stage", P), ("box", I * 6), ("x0", I), ("y0", I), ("z0", I), ("sx", I), ("sy", I),
        ("sz", I), ("kind", I), ("exact", I), ("axis", I), ("steps", I), ("nf", I),
        ("indices", P), ("weights", P), ("weights64", P), ("phasor", P), ("phasor64", P),
        ("flux", P), ("area", P),
    ]


class MonitorDesc(ctypes.Structure):
    """wfd_monitor (include/wfdtd.h)."""

    _fields_ = [
        ("indices", P), ("weights", P), ("phasor", P), ("flux", P), ("area", P),
        ("x0", I), ("y0", I), ("z0", I), ("sx", I), ("sy", I), ("sz", I),
        ("axis", I), ("exact", I), ("is_phasor", I), ("steps", I),
    ]


class GuideDesc(ctypes.Structure):
    """wfd_guide (include/wfdtd.h)."""

    _fields_ = [
        ("nx", I), ("ny", I), ("nz", I), ("spacing", F), ("dt", F),
        ("length", F), ("width", F), ("height", F),
        ("core_width", F), ("core_height", F),
        ("core_epsilon", F), ("clad_epsilon", F),
        ("sponge_width", F), ("sponge_rate", F),
        ("source_x", F), ("source_y", F), ("source_z", F),
        ("source_sigma_x", F), ("source_sigma_y", F), ("source_sigma_z", F),
        ("normal_average", I), ("layout_xyz", I),
    ]


_UPDATE = [P, P, P, P, L, P, I, P, L, P, L, P, P, I, I, I, I, I, F, I, P]

SIGNATURES = {
    "wfd_version": [],
    "wfd_last_error": [],
    "wfd_device_synchronize": [],
    "wfd_stream_synchronize": [P],
    "wfd_stream_create": [ctypes.POINTER(P)],
    "wfd_stream_destroy": [P],
    "wfd_graph_begin": [P],
    "wfd_graph_end": [P, ctypes.POINTER(P)],
    "wfd_graph_launch": [P, P],
    "wfd_graph_destroy": [P],
    "wfd_update_e": _UPDATE,
    "wfd_update_h": _UPDATE,
    "wfd_update_e_packed": _UPDATE,
    "wfd_update_h_packed": _UPDATE,
    "wfd_advance": [P, I, P],
    "wfd_yee_step": [ctypes.POINTER(YeePlan), P],
    "wfd_yee_pass": [ctypes.POINTER(YeePlan), ctypes.POINTER(YeeTargets), P],
    "wfd_batch_workspace_bytes": [I],
    "wfd_batch_prepare": [ctypes.POINTER(YeePlan), I, P, L, P],
    "wfd_batch_step": [ctypes.POINTER(YeePlan), I, P, P],
    "wfd_batch_pass_workspace_bytes": [I],
    "wfd_batch_pass_prepare": [ctypes.POINTER(YeePlan), ctypes.POINTER(YeeTargets), I, P, L, P],
    "wfd_batch_pass": [ctypes.POINTER(YeePlan), ctypes.POINTER(YeeTargets), I, P, P],
    "wfd_resident_layout_for": [ctypes.POINTER(YeePlan), I, I, I, ctypes.POINTER(ResidentLayout)],
    "wfd_resident_run": [ctypes.POINTER(YeePlan), ctypes.POINTER(ResidentLayout), P, I, I, P],
    "wfd_resident_run_noalloc": [ctypes.POINTER(YeePlan), ctypes.POINTER(ResidentLayout), I, I, P,
                                 P],
    "wfd_record": [P, P, P, P, P, P, P, P, P] + [I] * 13 + [P],
    "wfd_record_monitors": [P, P, P, P, ctypes.POINTER(MonitorDesc), I, I, I, I, P],
    "wfd_replay_monitors": [ctypes.POINTER(ReplayDesc), I, I, P, I, I, I, I, P],
    "wfd_dft_stage": [P, P, P, P, P] + [I] * 11 + [P],
    "wfd_dft_accumulate": [P, I, P, I, I, P, I, P, I, P],
    "wfd_record_volume": [P, P, P, P, P, P, P, I, I, I, I, I, P],
    "wfd_record_curl": [P, P, I, P] + [I] * 9 + [F, P],
    "wfd_adjoint_prepare": [P, P, P, P, I, P, I, I, I, F, I, P],
    "wfd_adjoint_gather": [P, P, P, I, P, P] + [I] * 10 + [P],
    "wfd_adjoint_gather_f64": [P, P, P, I, P, P] + [I] * 10 + [P],
    "wfd_adjoint_reverse_half": [P, P, P, P, P, P, I, P, I, P, P] + [I] * 9 + [F, I, P],
    "wfd_adjoint_reverse_half_f64": [P, P, P, P, P, P, I, P, I, P, P] + [I] * 9 + [F, I, P],
    "wfd_adjoint_inject": [P, P, P, P, P] + [I] * 11 + [P],
    "wfd_adjoint_inject_many": [P, I, P, P, P, I, I, I, I, P],
    "wfd_adjoint_inject_turn": [P, I, P, P, P, P, I, I, I, P],
    "wfd_write_design": [P, P] + [I] * 9 + [P],
    "wfd_driven_magnetic": [ctypes.POINTER(GuideDesc), P, P, P],
    "wfd_driven_electric": [ctypes.POINTER(GuideDesc), P, P, P, P, P],
    "wfd_observer_gather": [P, P, P, P, I, P, I, P],
    "wfd_count_nonfinite": [P, L, P, P],
    "wfd_record_mode": [P, P, P, P],
    "wfd_material_forward": [P, P, P, P, I, I, F, P],
    "wfd_material_pullback": [P, P, P, P, P, I, I, F, P],
    "wfd_gemm_bf16": [P, I, L, L, L, P, I, L, L, L, P, L, L, L, I, I, I, I, F, F, P],
    "wfd_to_bf16": [P, P, L, P],
    "wfd_round_bf16": [P, P, L, P],
    "wfd_bias_act": [P, P, P, P, L, I, I, P],
    "wfd_act_back": [P, P, P, L, I, P],
    "wfd_colsum_scratch": [L, I],
    "wfd_colsum": [P, L, I, P, P, L, F, P],
    "wfd_pool_mean_scratch": [I, I, I],
    "wfd_pool_mean": [P, I, I, I, P, P, L, P],
    "wfd_pool_mean_back": [P, I, I, I, P, F, P],
    "wfd_mse_partials": [L],
    "wfd_mse": [P, P, L, F, P, P, P],
    "wfd_adam": [P, P, P, P, P, L, F, F, F, F, I, F, P],
    "wfd_density_forward": [P, P] + [I] * 9 + [ctypes.c_double, ctypes.c_double, P],
    "wfd_density_pullback": [P, P, P, I, I, I, ctypes.c_double, ctypes.c_double, P],
}

_lib = None


def lib():
    """Load libwfdtd.so once; raises if it is absent or a symbol is missing."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc, gfx950)")
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, args in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = {"wfd_last_error": ctypes.c_char_p,
                          "wfd_batch_workspace_bytes": ctypes.c_int64,
                          "wfd_batch_pass_workspace_bytes": ctypes.c_int64,
                          "wfd_colsum_scratch": ctypes.c_int64,
                          "wfd_pool_mean_scratch": ctypes.c_int64,
                          "wfd_mse_partials": ctypes.c_int64}.get(name, ctypes.c_int)
        _lib = handle
    return _lib


def call(name, *args):
    """Invoke a C entry point and map its status to the reference's exceptions."""
    status = getattr(lib(), name)(*args)
    if status != WFD_OK:
        message = lib().wfd_last_error().decode(errors="replace")
        if status in (WFD_ERR_ARGUMENT, WFD_ERR_UNSUPPORTED):
            raise ValueError(f"{name}: {message}")
        raise RuntimeError(f"{name}: {message}")


def stream():
    """The current torch stream as a hipStream_t handle."""
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(a):
    """Device pointer of a tensor or Array (None passes as NULL)."""
    if a is None:
        return None
    t = getattr(a, "tensor", a)
    if not t.is_cuda:
        raise ValueError("libwfdtd arguments must be device tensors")
    if not t.is_contiguous():
        raise ValueError("libwfdtd arguments must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


class Handles:
    """The HIP graphs and streams one Python object (a solver, a batch, a
    driven guide) created through this library, destroyed by ``close()``.

    Every live set is also closed by an atexit hook (``shutdown``) that runs
    before the interpreter tears its objects down and before the HIP runtime
    (and any profiler attached to it) runs its own exit handlers: graph
    executables and streams left for the runtime's static destructors would
    otherwise be released after a profiling tool has already finalized.
    """

    _live = weakref.WeakValueDictionary()
    _order = itertools.count()

    def __init__(self):
        self._graphs = {}
        self._stream = None
        self.closed = False
        Handles._live[next(Handles._order)] = self

    def stream(self):
        """The owner's capture stream (created on first use)."""
        if self.closed:
            raise RuntimeError("library handles already closed")
        if self._stream is None:
            handle = ctypes.c_void_p()
            call("wfd_stream_create", ctypes.byref(handle))
            self._stream = handle
        return self._stream

    def begin(self):
        """Start capturing launches on the owner's stream."""
        s = self.stream()
        call("wfd_graph_begin", s)
        return s

    def end(self):
        """Finish the capture started by ``begin``; returns the graph handle."""
        graph = ctypes.c_void_p()
        call("wfd_graph_end", self._stream, ctypes.byref(graph))
        self._graphs[graph.value] = graph
        return graph

    def destroy(self, graph):
        """Destroy one graph of this owner (no-op for unknown handles)."""
        if graph is not None and self._graphs.pop(graph.value, None) is not None:
            call("wfd_graph_destroy", graph)

    def close(self):
        """Destroy every graph, then the stream; idempotent."""
        if self.closed:
            return
        self.closed = True
        graphs, self._graphs = list(self._graphs.values()), {}
        for g in graphs:
            call("wfd_graph_destroy", g)
        if self._stream is not None:
            s, self._stream = self._stream, None
            call("wfd_stream_destroy", s)


def release(handles, _finalizing=sys.is_finalizing):
    """``close()`` for an owner's ``__del__``: skipped while the interpreter
    is finalizing (the atexit hook has closed every live set by then) and
    silent on failure (a destructor must not raise)."""
    if handles is None or _finalizing():
        return
    try:
        handles.close()
    except Exception:
        pass


def shutdown():
    """Close every live handle set, newest first, after the device drained."""
    if _lib is None:
        return
    live = sorted(Handles._live.items(), key=lambda kv: -kv[0])
    if not live:
        return
    try:
        if torch.cuda.is_initialized():
            torch.cuda.synchronize()
    except Exception:
        pass
    for _, handles in live:
        try:
            handles.close()
        except Exception:
            pass


atexit.register(shutdown)
