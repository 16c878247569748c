"""Periodic vacuum Yee propagator on the libwfdtd kernels (drop-in for the Yee
branch of warp_fdtd/periodic.py:103-214).  It reuses update_e/update_h with
periodic=1, unit permittivity and inactive CPML rows, and records E_x of the
first cell every step.  The fourth-order RK4 branch of the reference is a
separate, non-Yee method and is not part of this package."""

import time

import numpy as np
import torch

from . import kernels as k
from ._lib import Handles, call, ptr, release, stream
from .array import Array


class Periodic:
    def __init__(self, n, cfl, method, shape=None, device="cuda:0"):
        if method not in ("yee", "rk4") or n < 4 or not 0 < cfl <= 0.99:
            raise ValueError("Require method yee/rk4, >=4 cells and 0<CFL<=0.99")
        if method == "rk4":
            raise ValueError("The RK4 propagator is not part of the uniform Yee path")
        self.shape = shape or (4, 4, n)
        self.ncells = int(np.prod(self.shape))
        self.n = n
        self.cfl = cfl
        self.method = method
        self.courant = cfl / np.sqrt(3)
        self.dt = 1550e-9 / n * self.courant / 299792458.0
        self.steps = round(20 * (1550e-9 / 299792458.0) / self.dt)
        self.device = torch.device(device)
        initial = np.zeros((6, *self.shape), np.float32)
        initial[0] = np.cos(2 * np.pi * np.arange(self.shape[2], dtype=np.float32) / n)
        self.initial = initial.reshape(6, -1)
        self.state = Array.from_numpy(self.initial, np.float32, self.device)
        self.E = self.state[:3]
        self.H = self.state[3:]
        self.step = Array.zeros(1, np.int32, self.device)
        self.signal = Array.zeros(self.steps, device=self.device)
        self.eps = Array(torch.ones(self.ncells, device=self.device))
        self.psi = Array.zeros((1, 1), device=self.device)
        coeff = np.zeros((3, max(self.shape), 3), np.float32)
        coeff[:, :, 1:] = 1
        self.coeff = Array.from_numpy(coeff, np.float32, self.device)
        self.dummy = Array.zeros((3, 1), device=self.device)
        self.graphs = {}
        self._handles = Handles()

    def launch_step(self):
        """periodic.py:136-188 (Yee branch)."""
        args = [self.eps, self.psi, self.coeff, self.dummy, self.dummy, self.step, self.dummy,
                self.steps, *self.shape, -1, self.courant, 1]
        k.update_e(self.E, self.H, *args)
        k.update_h(self.H, self.E, *args)
        call("wfd_record_mode", ptr(self.E), ptr(self.signal), ptr(self.step), stream())

    def reset(self):
        self.state.assign(self.initial)
        self.step.zero_()
        self.signal.zero_()
        torch.cuda.synchronize(self.device)

    def compile(self):
        start = time.perf_counter()
        self.launch_step()
        torch.cuda.synchronize(self.device)
        self.reset()
        for g in self.graphs.values():
            self._handles.destroy(g)
        self.graphs = {}
        for length in {64, self.steps % 64} - {0}:
            s = self._handles.begin()
            ext = torch.cuda.ExternalStream(s.value, device=self.device)
            try:
                with torch.cuda.stream(ext):
                    for _ in range(length):
                        self.launch_step()
            finally:
                graph = self._handles.end()
            self.graphs[length] = graph
        self.reset()
        return time.perf_counter() - start

    def run(self):
        s = stream()
        for _ in range(self.steps // 64):
            call("wfd_graph_launch", self.graphs[64], s)
        if self.steps % 64:
            call("wfd_graph_launch", self.graphs[self.steps % 64], s)
        torch.cuda.synchronize(self.device)

    def close(self):
        """Destroy the captured graphs and stream (idempotent)."""
        self.graphs = {}
        self._handles.close()

    def __del__(self, _release=release):
        # (bound at definition: module globals may already be None when an
        # object is collected during interpreter shutdown)
        _release(getattr(self, "_handles", None))
