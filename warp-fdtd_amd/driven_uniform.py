"""Direct structured Yee solver for the driven dielectric guide on libwfdtd
(drop-in for warp_fdtd/driven_uniform.py).

Staggered per-component arrays (E_c of shape n+1 except along c, H_c of
shape n except n+1 along c), a cubic conductivity sponge, a Gaussian impressed
current on E_y and a PEC exterior.  H half-step first, then E, then the
counter advance; ``compile`` captures a 32-step HIP graph.
"""

import ctypes
import time

import numpy as np
import torch

from ._lib import GuideDesc, Handles, call, ptr, release, stream
from .array import Array


class DrivenUniform:
    """driven_uniform.py:277-406"""

    def __init__(self, problem, spacing, steps, cfl=0.7, device="cuda:0", layout="xyz",
                 normal_average=False):
        if layout not in ("xyz", "flat"):
            raise ValueError("layout must be xyz or flat")
        self.layout = layout
        self.normal_average = int(normal_average)
        self.problem = problem
        self.hspacing = spacing
        self.shape = tuple(int(round(x / spacing)) for x in problem.lengths)
        if any(abs(n * spacing - x) > 1e-10 for n, x in zip(self.shape, problem.lengths)):
            raise ValueError("Spacing must divide domain")
        self.n = self.shape
        self.device = torch.device(device)
        self.dt = cfl * spacing * np.sqrt(problem.clad_epsilon) / np.sqrt(3)
        self.eshapes = [tuple(n if a == c else n + 1 for a, n in enumerate(self.shape))
                        for c in range(3)]
        self.hshapes = [tuple(n + 1 if a == c else n for a, n in enumerate(self.shape))
                        for c in range(3)]
        self.e = Array.zeros(sum(int(np.prod(s)) for s in self.eshapes), device=self.device)
        self.h = Array.zeros(sum(int(np.prod(s)) for s in self.hshapes), device=self.device)
        self.counter = Array.zeros(1, np.int32, self.device)
        self.wave = Array.from_numpy(
            problem.waveform((np.arange(steps + 64) + 0.5) * self.dt).astype(np.float32),
            np.float32, self.device)
        self.arrays = [self.e, self.h, self.counter, self.wave]
        self.graph = None
        self.graph_steps = 32
        self._handles = Handles()
        p = problem
        g = GuideDesc()
        g.nx, g.ny, g.nz = self.shape
        g.spacing, g.dt = spacing, self.dt
        g.length, g.width, g.height = p.lengths
        g.core_width, g.core_height = p.core_width, p.core_height
        g.core_epsilon, g.clad_epsilon = p.core_epsilon, p.clad_epsilon
        g.sponge_width, g.sponge_rate = p.sponge_width, p.sponge_rate
        g.source_x, g.source_y, g.source_z = p.source_x, p.width / 2, p.height / 2
        g.source_sigma_x = p.source_sigma_x
        g.source_sigma_y = g.source_sigma_z = p.source_sigma_transverse
        g.normal_average = self.normal_average
        g.layout_xyz = int(layout == "xyz")
        self._desc = g

    @property
    def owned_bytes(self):
        return sum(a.capacity for a in self.arrays)

    def reset(self):
        self.e.zero_()
        self.h.zero_()
        self.counter.zero_()

    def magnetic_step(self):
        call("wfd_driven_magnetic", ctypes.byref(self._desc), ptr(self.e), ptr(self.h), stream())

    def step(self):
        """driven_uniform.py:354-384 — magnetic, electric, advance."""
        self.magnetic_step()
        call("wfd_driven_electric", ctypes.byref(self._desc), ptr(self.h), ptr(self.e),
             ptr(self.wave), ptr(self.counter), stream())
        call("wfd_advance", ptr(self.counter), 1, stream())

    def compile(self):
        start = time.perf_counter()
        self.step()
        torch.cuda.synchronize(self.device)
        self.reset()
        torch.cuda.synchronize(self.device)
        s = self._handles.begin()
        ext = torch.cuda.ExternalStream(s.value, device=self.device)
        try:
            with torch.cuda.stream(ext):
                for _ in range(self.graph_steps):
                    self.step()
        finally:
            graph = self._handles.end()
        self._handles.destroy(self.graph)
        self.graph = graph
        self.reset()
        return time.perf_counter() - start

    def run(self, steps):
        """driven_uniform.py:397-406 — graphs of 32 steps, then an eager tail."""
        if self.graph is None:
            for _ in range(steps):
                self.step()
        else:
            s = stream()
            for _ in range(steps // self.graph_steps):
                call("wfd_graph_launch", self.graph, s)
            for _ in range(steps % self.graph_steps):
                self.step()
        torch.cuda.synchronize(self.device)

    def close(self):
        """Destroy the captured graph and stream (idempotent)."""
        self.graph = None
        self._handles.close()

    def __del__(self, _release=release):
        # (bound at definition: module globals may already be None when an
        # object is collected during interpreter shutdown)
        _release(getattr(self, "_handles", None))
