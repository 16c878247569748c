"""Common-physical-point Yee sampling with explicit H time centering for the
DrivenUniform solver (drop-in for warp_fdtd/guide_observer.py; the octree
branch belongs to the adaptive-mesh prototype)."""

from itertools import product

import numpy as np
import torch

from ._lib import call, ptr, stream
from .array import Array


def points(problem):
    """guide_observer.py:28-34"""
    axes = (
        np.arange(0.45, problem.length - 0.44, 0.1),
        np.arange(problem.width / 2 - 0.775, problem.width / 2 + 0.776, 0.1),
        np.arange(problem.height / 2 - 0.775, problem.height / 2 + 0.776, 0.1),
    )
    return axes, np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 3)


def _stencils(xyz, h, shape, component, electric, offset):
    """Trilinear corner indices and weights of one staggered component grid
    (guide_observer.py:37-49) for many points at once."""
    staggering = np.zeros(3) if electric else np.full(3, 0.5)
    staggering[component] = 0.5 if electric else 0
    index = xyz / h - staggering
    base = np.clip(np.floor(index).astype(int), 0, np.array(shape) - 2)
    fraction = index - base
    indices = np.empty((len(xyz), 8), np.int64)
    weights = np.empty((len(xyz), 8), np.float64)
    for j, corner in enumerate(product((0, 1), repeat=3)):
        p = base + np.array(corner)
        indices[:, j] = offset + np.ravel_multi_index(tuple(p.T), shape)
        weights[:, j] = np.prod(np.where(np.array(corner, bool), fraction, 1 - fraction), axis=1)
    return indices, weights


class GuideObserver:
    """guide_observer.py:52-131 (uniform branch)."""

    def __init__(self, solver, problem):
        self.solver = solver
        self.axes, self.xyz = points(problem)
        n = len(self.xyz)
        indices = [np.empty((3 * n, 8), np.int32) for _ in range(2)]
        weights = [np.empty((3 * n, 8), np.float32) for _ in range(2)]
        for electric, shapes in ((True, solver.eshapes), (False, solver.hshapes)):
            offsets = np.cumsum([0] + [int(np.prod(s)) for s in shapes])
            for c in range(3):
                ix, w = _stencils(self.xyz, solver.hspacing, shapes[c], c, electric,
                                  int(offsets[c]))
                indices[not electric][c * n:(c + 1) * n] = ix
                weights[not electric][c * n:(c + 1) * n] = w
        self.indices = [Array.from_numpy(v, np.int32, solver.device) for v in indices]
        self.weights = [Array.from_numpy(v, np.float32, solver.device) for v in weights]
        self.out = [Array.zeros(3 * n, device=solver.device) for _ in range(2)]

    def sample(self):
        """Clone H, one magnetic half-step for time centring, gather E and
        0.5*(H_new + H_old), restore H (guide_observer.py:103-131)."""
        s = self.solver
        old = Array(s.h.tensor.clone())
        s.magnetic_step()
        for c, field in enumerate((s.e, s.h)):
            call("wfd_observer_gather", ptr(self.indices[c]), ptr(self.weights[c]), ptr(field),
                 ptr(old), c, ptr(self.out[c]), len(self.out[c]), stream())
        s.h.tensor.copy_(old.tensor)
        torch.cuda.synchronize(s.device)
        return np.concatenate([v.numpy().reshape(3, -1) for v in self.out])
