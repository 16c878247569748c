"""Design-parameter material map on libwfdtd (drop-in for warp_fdtd/material.py).

19x19 Gaussian filter (sigma = 3 cells, edge clamp) and the tanh subpixel
projection to inv_eps = 1/(1 + 11.25 value); ``pullback`` applies the
hand-derived transpose of both maps (the reference records them on wp.Tape).
"""

import numpy as np
import torch

from ._lib import call, ptr, stream
from .array import Array


def filter_weights():
    """material.py:73-75 — normalised 19x19 Gaussian, sigma = 3 cells."""
    xy = np.arange(-9, 10, dtype=np.float32)
    weights = np.exp(-(xy[:, None] ** 2 + xy[None, :] ** 2) / 18).astype(np.float32)
    return weights / weights.sum()


class Material:
    """material.py:58-102"""

    def __init__(self, params, beta, device="cuda:0"):
        params = np.asarray(params)
        if (params.ndim != 2 or min(params.shape) < 2 or not np.isfinite(params).all()
                or not np.isfinite(beta) or beta <= 0):
            raise ValueError("Require finite 2D parameters and positive finite beta")
        self.device = torch.device(device)
        self.p = Array.from_numpy(params.astype(np.float32), np.float32, self.device)
        self.filtered = Array.zeros(params.shape, device=self.device)
        self.inv_eps = Array.zeros(params.shape, device=self.device)
        self.weights = Array.from_numpy(filter_weights(), np.float32, self.device)
        self.beta = float(beta)
        self._scratch = Array.zeros((4, *params.shape), device=self.device)
        call("wfd_material_forward", ptr(self.p), ptr(self.weights), ptr(self.filtered),
             ptr(self.inv_eps), *params.shape, float(self.beta), stream())

    def pullback(self, gradient):
        """d(sum(gradient * inv_eps))/d(params) as a NumPy array."""
        g = Array.from_numpy(np.asarray(gradient, np.float32).reshape(self.p.shape),
                             np.float32, self.device)
        out = Array.zeros(self.p.shape, device=self.device)
        call("wfd_material_pullback", ptr(self.weights), ptr(self.filtered), ptr(g), ptr(out),
             ptr(self._scratch), *self.p.shape, float(self.beta), stream())
        return out.numpy()
