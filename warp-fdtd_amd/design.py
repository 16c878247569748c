"""Density-parameterised inverse design on libwfdtd.

The design is a density rho in [0, 1] on an (sx, sy) grid of in-plane voxels;
the linear map eps = eps_lo + rho (eps_hi - eps_lo) fills the design box's
z-layers (wfd_density_forward).  ``TransmissionProblem`` owns a checkpointed
``Adjoint`` of a density bundle (cases.bend_case) and evaluates the modal
transmission T at one wavelength and dT/drho: the adjoint accumulates
d(-T)/d(inv_eps) over the box cells in float64, and wfd_density_pullback
chains it through the map.  ``transmission(rho, problem)`` exposes the pair
as a torch.autograd.Function.

No reference counterpart: the reference's design path is the 80x80 filtered
Material of adjoint.py:441-469 (kept in material.py / Adjoint.differentiate).
"""

import time

import numpy as np
import torch

from ._lib import call, ptr, stream
from .adjoint import Adjoint

EPS_OXIDE = 2.085
EPS_SI = 12.08


class DensityMap:
    """eps = eps_lo + rho (eps_hi - eps_lo) through the box (x0, y0, z0, sx, sy, sz)."""

    def __init__(self, box, eps_lo=EPS_OXIDE, eps_hi=EPS_SI):
        self.box = tuple(int(v) for v in box)
        if len(self.box) != 6 or min(self.box[3:]) < 1:
            raise ValueError("box is (x0, y0, z0, sx, sy, sz) with positive sizes")
        if not (eps_lo > 0 and eps_hi > 0):
            raise ValueError("Permittivities must be positive")
        self.eps_lo, self.eps_hi = float(eps_lo), float(eps_hi)
        self.plane = self.box[3:5]

    def _rho(self, rho, device):
        t = torch.as_tensor(rho, dtype=torch.float64, device=device)
        if tuple(t.shape) != self.plane:
            raise ValueError(f"rho must have shape {self.plane}")
        if not torch.isfinite(t).all() or t.min() < 0 or t.max() > 1:
            raise ValueError("rho must lie in [0, 1]")
        return t.contiguous()

    def write(self, rho, solver):
        """Write 1/eps of the design into solver.eps (in place)."""
        r = self._rho(rho, solver.device)
        call("wfd_density_forward", ptr(r), ptr(solver.eps), *solver.shape, *self.box,
             self.eps_lo, self.eps_hi, stream())
        return r

    def pullback(self, rho, grad_inv_eps):
        """d/d(rho) of a scalar from its float64 d/d(inv_eps) over the box cells."""
        g = getattr(grad_inv_eps, "tensor", grad_inv_eps)
        if g.dtype != torch.float64 or g.numel() != int(np.prod(self.box[3:])):
            raise ValueError("grad_inv_eps must be float64 over the design box")
        r = self._rho(rho, g.device)
        g = g.contiguous()
        out = torch.empty(self.plane, dtype=torch.float64, device=g.device)
        call("wfd_density_pullback", ptr(r), ptr(g), ptr(out), *self.box[3:],
             self.eps_lo, self.eps_hi, stream())
        return out


class TransmissionProblem:
    """Modal transmission T(rho) of a density bundle and its adjoint gradient.

    ``case`` is a bundle written by cases.bend_case (its settings hold the
    design box); ``wavelength`` picks the DFT wavelength of the objective
    (optional when the planes hold one); ``checkpoint_every`` is the adjoint's
    checkpoint interval (None keeps the full curl history).
    """

    def __init__(self, case, wavelength=None, checkpoint_every=64, device="cuda:0", **solver_kw):
        from . import detectors

        meta_box = _design_box(case)
        self.adjoint = Adjoint(case, checkpoint_every=checkpoint_every, device=device,
                               box=meta_box, require_reference_bend=False,
                               objective_wavelength=wavelength, **solver_kw)
        detectors.attach(self.adjoint)
        self.map = DensityMap(meta_box)
        self._forward_rho = None
        self._compiled = False
        self.timings = {}

    def _compile(self):
        if not self._compiled:
            self.adjoint.compile()
            self._compiled = True

    def forward(self, rho):
        """T at the objective wavelength for the density rho (sx, sy)."""
        self._compile()
        a = self.adjoint
        t0 = time.perf_counter()
        r = self.map.write(rho, a)
        a.reset()
        a.run()
        self._forward_rho = r.clone()
        self.timings["forward_s"] = time.perf_counter() - t0
        return a.transmission()

    def backward(self, rho):
        """dT/drho (sx, sy) float64 tensor for the design of the last forward
        (re-runs the forward when rho differs from it)."""
        r = self.map._rho(rho, self.adjoint.device)
        if self._forward_rho is None or not torch.equal(r, self._forward_rho):
            self.forward(r)
        t0 = time.perf_counter()
        self._prepare_backward()
        self.adjoint.backward()
        out = self._finish_backward(r)
        self.timings["backward_s"] = time.perf_counter() - t0
        return out

    def _prepare_backward(self):
        """The reverse's setup after a forward: adjoint sources, reverse
        graphs, zeroed reverse state."""
        a = self.adjoint
        a.prepare_sources()
        if not a.reverse_graphs:
            # the reverse graphs capture the cotangent buffers, which exist
            # once prepare_sources has run; later calls refresh them in place
            a.compile_reverse()
        a.reset_reverse()

    def _finish_backward(self, r):
        """dT/drho from the finished reverse (the adjoint accumulates
        d(-T)/d(inv_eps))."""
        out = -self.map.pullback(r, self.adjoint.grad)
        torch.cuda.synchronize(self.adjoint.device)
        return out

    def value_and_grad(self, rho):
        T = self.forward(rho)
        return T, self.backward(rho)


def _design_box(case):
    from .solver import load_case

    _, meta, _ = load_case(case)
    box = meta["settings"].get("design_box")
    if box is None:
        raise ValueError("The bundle declares no design_box (write it with cases.bend_case)")
    return tuple(int(v) for v in box)


class Transmission(torch.autograd.Function):
    """T(rho) with dT/drho from the adjoint."""

    @staticmethod
    def forward(ctx, rho, problem):
        ctx.problem = problem
        ctx.save_for_backward(rho)
        value = problem.forward(rho.detach())
        return torch.tensor(value, dtype=torch.float64, device=rho.device)

    @staticmethod
    def backward(ctx, grad_output):
        (rho,) = ctx.saved_tensors
        grad = ctx.problem.backward(rho.detach())
        return grad_output * grad.to(rho.dtype), None


def transmission(rho, problem):
    """Differentiable T(rho) (torch.autograd)."""
    return Transmission.apply(rho, problem)


class BatchTransmission:
    """Forward + adjoint of many density designs of uneven grids.

    The forwards of every design run together: one ``Batch`` graph per chunk
    steps all grids, and each design's field checkpoint is saved between
    chunks every ``checkpoint_every`` steps (a multiple of ``graph_steps``).
    Each design's reverse then runs on its own recompute/reverse graphs.
    With ``checkpoint_every`` None the batched forward records every design's
    whole design-box curl history instead (steps x 3 x box cells floats per
    design) and the reverses recompute nothing.  ``single_pass`` is the
    batch's (``Batch``).  Values and gradients are bitwise those of the
    designs' own ``TransmissionProblem``s.
    """

    def __init__(self, cases, wavelength=None, checkpoint_every=192, graph_steps=64,
                 device="cuda:0", single_pass="auto"):
        from .batch import Batch

        if checkpoint_every is not None and checkpoint_every % graph_steps:
            raise ValueError("checkpoint_every must be a multiple of graph_steps (or None)")
        self.problems = [TransmissionProblem(c, wavelength=wavelength,
                                             checkpoint_every=checkpoint_every, device=device,
                                             graph_steps=graph_steps, resident=False)
                         for c in cases]
        self.batch = Batch([p.adjoint for p in self.problems], graph_steps=graph_steps,
                           single_pass=single_pass)
        self.checkpoint_every = checkpoint_every
        self.graph_steps = graph_steps
        self.cells = self.batch.cells
        self.steps = self.batch.steps
        self._compiled = False
        self.timings = {}
        self.log = lambda message: None

    def compile(self):
        if not self._compiled:
            for k, p in enumerate(self.problems):
                p._compile()
                if k % 8 == 7:
                    self.log(f"compiled {k + 1}/{len(self.problems)} designs")
            self.batch.compile()
            self._compiled = True

    def forward(self, rhos):
        """T of every design (list of floats) for the densities ``rhos``."""
        if len(rhos) != len(self.problems):
            raise ValueError("one density per design")
        self.compile()
        t0 = time.perf_counter()
        written = [p.map.write(r, p.adjoint) for p, r in zip(self.problems, rhos)]
        self.batch.reset()
        active = np.zeros(self.steps, bool)
        for p in self.problems:
            active |= p.adjoint._windows()
        K, G = self.checkpoint_every or self.steps, self.graph_steps
        s = stream()
        for c, start in enumerate(range(0, self.steps, K)):
            if self.checkpoint_every is not None:
                for p in self.problems:
                    for dst, src in zip(p.adjoint.checkpoints[c], p.adjoint._state()):
                        dst.tensor.copy_(src.tensor)
            t, end = start, min(start + K, self.steps)
            while t < end:
                length = min(G, end - t)
                g = (self.batch.graphs if active[t:t + length].any()
                     else self.batch.quiet_graphs)[length]
                call("wfd_graph_launch", g, s)
                t += length
        torch.cuda.synchronize(self.problems[0].adjoint.device)
        self.timings["forward_s"] = time.perf_counter() - t0
        for p, r in zip(self.problems, written):
            p._forward_rho = r.clone()
        return [p.adjoint.transmission() for p in self.problems]

    def backward(self, streams=2):
        """dT/drho of every design for the densities of the last forward.
        Every design's reverse is set up first; the reverses then run on
        ``streams`` streams in turn, so that one design's latency-bound
        launches (the inject turn) overlap another's reverse halves (the
        designs share no buffers)."""
        t0 = time.perf_counter()
        for k, p in enumerate(self.problems):
            p._prepare_backward()
            if k % 8 == 7:
                self.log(f"adjoint set up {k + 1}/{len(self.problems)}")
        lanes = [torch.cuda.Stream(device=self.problems[0].adjoint.device)
                 for _ in range(max(1, int(streams)))]
        torch.cuda.synchronize(self.problems[0].adjoint.device)
        for k, p in enumerate(self.problems):
            with torch.cuda.stream(lanes[k % len(lanes)]):
                p.adjoint.launch_backward()
        torch.cuda.synchronize(self.problems[0].adjoint.device)
        grads = [p._finish_backward(p._forward_rho) for p in self.problems]
        self.timings["backward_s"] = time.perf_counter() - t0
        return grads
