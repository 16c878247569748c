"""Surrogate-assisted inverse design of the density bend (no reference counterpart).

* ``surrogate_case`` writes the bend bundle the surrogate learns: the
  200x200x80 bend at 20 nm with a 100x100-voxel (2x2 um) design region,
  10-cell CPML, 4000 steps, modal DFT planes at 8 wavelengths (1.50-1.60 um)
  plus the objective's 1.55 um.
* ``Evaluator`` runs the solver: forward + checkpointed adjoint labels (T
  spectrum and dT/drho at 1.55 um), single forward verifications and single
  adjoint steps.
* ``train`` fits the FNO (fno.py) to (rho -> T spectrum, dT/drho) on the
  library's own BF16 kernels.
* ``optimize`` runs gradient ascent on T at the objective wavelength either
  with the adjoint at every step (``surrogate=None``) or with surrogate
  proposals, each re-verified by a full FDTD forward and replaced by an
  adjoint step when the surrogate's transmission error exceeds ``tolerance``.
"""

import time

import numpy as np
import torch

from . import detectors
from .cases import bend_case, design_density
from .design import DensityMap, TransmissionProblem, _design_box
from .solver import Solver

SPECTRUM = np.linspace(1.50e-6, 1.60e-6, 8)
OBJECTIVE = 1.55e-6
# the DFT planes record the spectrum and then the objective's wavelength
WAVELENGTHS = np.append(SPECTRUM, OBJECTIVE)


def surrogate_case(path, side=100, shape=(200, 200, 80), spacing=20e-9, steps=4000, pml=10,
                   pulse_duration=60e-15):
    """The 90-degree bend with a side x side voxel design region (one voxel per
    in-plane cell) and modal DFT planes at WAVELENGTHS (cases.bend_case)."""
    return bend_case(path, np.full((side, side), 0.5), shape=shape, spacing=spacing,
                     steps=steps, pml=pml, design=(side, side), wavelengths=list(WAVELENGTHS),
                     wavelength=OBJECTIVE, pulse_duration=pulse_duration)


def objective_index():
    """Index of the objective's wavelength (1.55 um) in a label's spectrum."""
    return len(SPECTRUM)


class Evaluator:
    """FDTD evaluations of densities on one surrogate bundle: labels and
    adjoint steps by forward + checkpointed adjoint (design.TransmissionProblem),
    verifications by a plain forward (Solver, no checkpoints)."""

    def __init__(self, case, checkpoint_every=64, device="cuda:0", log=None):
        self.log = log or (lambda message: None)
        self.case = case
        self.problem = TransmissionProblem(case, wavelength=OBJECTIVE,
                                           checkpoint_every=checkpoint_every, device=device)
        self.checker = Solver(case, device=device)
        detectors.attach(self.checker)
        self.checker.compile()
        self.map = DensityMap(_design_box(case))
        self.side = self.map.plane[0]
        self.counts = {"forward": 0, "adjoint": 0, "labels": 0}

    def adjoint(self, rho):
        """(T spectrum (9,), dT/drho at the objective wavelength) of one design."""
        _, g = self.problem.value_and_grad(rho)
        self.counts["adjoint"] += 1
        return self.problem.adjoint.transmission_spectrum(), g.cpu().numpy()

    def label(self, rhos, every=64):
        """(T spectra (n, 9), dT/drho (n, side, side)) of the densities."""
        spectra, grads = [], []
        for k, rho in enumerate(rhos):
            spec, g = self.adjoint(rho)
            spectra.append(spec)
            grads.append(g)
            if (k + 1) % every == 0:
                self.log(f"labelled {k + 1}/{len(rhos)} designs")
        self.counts["labels"] += len(spectra)
        return np.array(spectra), np.array(grads)

    def transmission(self, rho):
        """T spectrum (9,) by one full FDTD forward (the verification)."""
        s = self.checker
        self.map.write(rho, s)
        s.reset()
        s.run()
        self.counts["forward"] += 1
        return detectors.dft_transmission(s)


def design(seed, index, side):
    """Blurred-random density number ``index`` of the seeded family (sigma 2
    voxels), independent of every other index."""
    return design_density(np.random.default_rng([int(seed), int(index)]), (side, side),
                          blurred=True)


def random_designs(count, side, seed=0, start=0):
    """Designs start .. start + count - 1 of the seeded family."""
    return [design(seed, k, side) for k in range(start, start + count)]


class Normalizer:
    def __init__(self, T, G):
        self.t_mean = T.mean(0)
        self.t_std = T.std(0) + 1e-6
        self.g_scale = float(np.sqrt((G ** 2).mean())) + 1e-12


class _Fitter:
    """Adam steps of the FNO on (rho -> normalised T spectrum, dT/drho)
    batches held on the GPU (MSE on both heads, the library's kernels)."""

    def __init__(self, net, norm, grad_weight=1.0):
        from ._lib import lib

        self.net, self.norm, self.grad_weight = net, norm, grad_weight
        B, dev = net.B, net.device
        self.d_t = torch.zeros(B, net.F, device=dev)
        self.d_g = torch.zeros(B, net.S, net.S, device=dev)
        self.part_t = torch.zeros(int(lib().wfd_mse_partials(B * net.F)), device=dev)
        self.part_g = torch.zeros(int(lib().wfd_mse_partials(B * net.S * net.S)), device=dev)

    def tensors(self, rhos, T, G):
        """Device tensors of densities and normalised labels."""
        dev, norm = self.net.device, self.norm
        x = torch.from_numpy(np.asarray(rhos, np.float32)).to(dev)
        tn = torch.from_numpy(((np.asarray(T, np.float64) - norm.t_mean) / norm.t_std)
                              .astype(np.float32)).to(dev)
        gn = torch.from_numpy((np.asarray(G, np.float64) / norm.g_scale).astype(np.float32)).to(dev)
        return x, tn, gn

    def step(self, x, tn, gn, idx, rate):
        """One Adam step on the rows idx (a device index tensor of net.B rows);
        returns the batch loss as a device scalar."""
        from ._lib import call, stream

        net, B = self.net, self.net.B
        t_out, g_out = net.forward(x.index_select(0, idx).contiguous())
        tt = tn.index_select(0, idx).contiguous()
        gt = gn.index_select(0, idx).contiguous()
        call("wfd_mse", t_out.data_ptr(), tt.data_ptr(), B * net.F, 2.0 / (B * net.F),
             self.d_t.data_ptr(), self.part_t.data_ptr(), stream())
        call("wfd_mse", g_out.data_ptr(), gt.data_ptr(), B * net.S * net.S,
             2.0 * self.grad_weight / (B * net.S * net.S), self.d_g.data_ptr(),
             self.part_g.data_ptr(), stream())
        net.backward(self.d_t, self.d_g)
        net.adam(rate)
        return (self.part_t.sum() / (B * net.F) + self.grad_weight * self.part_g.sum()
                / (B * net.S * net.S))


def train(net, rhos, T, G, epochs=40, lr=2e-3, grad_weight=1.0, seed=0, norm=None, log=None):
    """Fit the FNO to (rho -> T spectrum, dT/drho) with Adam on the GPU.
    Returns the normalizer and the per-epoch mean loss."""
    norm = norm or Normalizer(np.asarray(T, np.float64), np.asarray(G, np.float64))
    fit = _Fitter(net, norm, grad_weight)
    x, tn, gn = fit.tensors(rhos, T, G)
    B = net.B
    rng = np.random.default_rng(seed)
    history = []
    n = len(rhos)
    for epoch in range(epochs):
        order = rng.permutation(n)
        total, batches = 0.0, 0
        rate = lr * 0.5 * (1 + np.cos(np.pi * epoch / max(epochs, 1)))
        for start in range(0, n - B + 1, B):
            idx = torch.from_numpy(order[start:start + B]).to(net.device)
            total += float(fit.step(x, tn, gn, idx, rate))
            batches += 1
        history.append(total / max(batches, 1))
        if log:
            log(f"epoch {epoch} loss {history[-1]:.5f}")
    return norm, history


class OnlineTuner:
    """Fine-tuning during the design loop: every adjoint label the loop
    computes (a fallback) joins a buffer, and ``iters`` Adam steps run on
    batches of the newest ``recent`` labels topped up with random training
    designs (so the network adapts to the trajectory without forgetting)."""

    def __init__(self, net, norm, data, iters=8, lr=3e-4, recent=8, seed=0):
        self.fit = _Fitter(net, norm)
        self.x, self.tn, self.gn = self.fit.tensors(*data)
        self.iters, self.lr, self.recent = iters, lr, recent
        self.rng = np.random.default_rng(seed)
        self.buf = []

    def add(self, rho, spec, grad):
        self.buf.append((np.asarray(rho), np.asarray(spec), np.asarray(grad)))
        net = self.fit.net
        new = self.buf[-self.recent:]
        bx, bt, bg = self.fit.tensors([b[0] for b in new], [b[1] for b in new],
                                      [b[2] for b in new])
        x = torch.cat([bx, self.x])
        tn = torch.cat([bt, self.tn])
        gn = torch.cat([bg, self.gn])
        k = len(new)
        for _ in range(self.iters):
            fill = self.rng.choice(len(self.x), net.B - k, replace=False) + k
            idx = torch.from_numpy(np.concatenate([np.arange(k), fill])).to(net.device)
            self.fit.step(x, tn, gn, idx, self.lr)


def inference_copy(net):
    """A batch-1 FNO holding the trained net's parameters, for the loop's
    one-design predictions (the training net evaluates net.B designs)."""
    from .fno import FNO

    one = FNO(net.S, width=net.W, modes=net.m, hidden=net.H, outputs=net.F, layers=net.L,
              batch=1, device=net.device)
    one.params.copy_(net.params)
    one.sync_bf16()
    return one


def predict(net, norm, rho):
    """(T spectrum (8,), dT/drho (side, side)) of one density."""
    x = torch.as_tensor(np.asarray(rho, np.float32), device=net.device)
    t, g = net.forward(x[None].expand(net.B, net.S, net.S).contiguous())
    t = t[0].detach().cpu().double().numpy() * norm.t_std + norm.t_mean
    g = g[0].detach().cpu().double().numpy() * norm.g_scale
    return t, g


def _ascent(rho, g, step):
    scale = np.abs(g).max()
    if not np.isfinite(scale) or scale == 0:
        return rho
    return np.clip(rho + step * g / scale, 0.0, 1.0)


def optimize(evaluator, rho0, steps=200, step=0.01, surrogate=None, tolerance=0.01, log=None,
             inner=2, max_inner=16, lagged=False, target=None, tuner=None):
    """Gradient ascent on T at the objective wavelength (largest density
    change ``step`` per ascent step).

    ``surrogate=None``: the plain loop, an adjoint (forward + checkpointed
    reverse) at every step, ``steps`` steps, then a forward of the final
    design.

    ``surrogate=(net, norm)``: each iteration proposes m ascent steps (m
    starts at ``inner``) along the anchored model gradient g_a + g_net(rho) -
    g_net(rho_a), where g_a is the adjoint gradient of the anchor design
    rho_a; the surrogate's T of the proposal is the last verified T plus the
    trapezoid-rule integral of that gradient along the m steps.  A full FDTD
    forward verifies every proposal: within ``tolerance`` (relative) of the
    prediction and above the current T, it is accepted and m doubles (up to
    ``max_inner``); otherwise m halves and, unless the anchor already is the
    current design, the adjoint of the current design becomes the new anchor
    (the fallback).  A one-step proposal from a fresh anchor is a plain
    adjoint step and is always accepted.  ``lagged`` drops
    the network's correction (g_a alone: the method without the network).
    ``target``: stop once the verified T reaches it.  ``tuner`` (OnlineTuner)
    learns from every adjoint label the loop computes.

    Returns a dict: verified T per iteration ("history", with the wall time of
    each in "times"), the final design and T, the relative error of every
    surrogate prediction, fallbacks, evaluation counts and wall time."""
    f = objective_index()
    rho = np.array(rho0, np.float64)
    start = time.perf_counter()
    history, times, errors, fallbacks, accepted, sizes = [], [], [], 0, 0, []
    counts0 = dict(evaluator.counts)

    def now():
        return time.perf_counter() - start

    if surrogate is None:
        for k in range(steps):
            spec, g = evaluator.adjoint(rho)
            history.append(float(spec[f]))
            times.append(now())
            rho = _ascent(rho, g, step)
            if log and k % 20 == 0:
                log(f"adjoint step {k}: T = {history[-1]:.5f}")
        final = float(evaluator.transmission(rho)[f])
        history.append(final)
        times.append(now())
    else:
        net, norm = surrogate

        def anchor_at(design):
            spec, g = evaluator.adjoint(design)
            if tuner is not None:
                tuner.add(design, spec, g)
            g_net = None if lagged else predict(net, norm, design)[1]
            return float(spec[f]), (np.array(design), g, g_net)

        current, anchor = anchor_at(rho)
        history.append(current)
        times.append(now())
        m = max(1, int(inner))
        for k in range(steps):
            if target is not None and current >= target:
                break
            fresh = np.array_equal(anchor[0], rho)

            def model_gradient(design):
                if lagged:
                    return anchor[1]
                return anchor[1] + predict(net, norm, design)[1] - anchor[2]

            path, g_here, t_hat = rho, anchor[1] if fresh else model_gradient(rho), current
            for _ in range(m):
                nxt = _ascent(path, g_here, step)
                g_next = model_gradient(nxt)
                t_hat += 0.5 * float(((g_here + g_next) * (nxt - path)).sum())
                path, g_here = nxt, g_next
            t_true = float(evaluator.transmission(path)[f])
            errors.append(abs(t_hat - t_true) / max(abs(t_true), 1e-12))
            sizes.append(m)
            if (errors[-1] <= tolerance and t_true > current) or (fresh and m == 1):
                rho, current = path, t_true
                accepted += 1
                m = min(2 * m, max_inner)
            else:
                m = max(1, m // 2)
                if not fresh:  # (an anchor at this design already holds its gradient)
                    fallbacks += 1
                    current, anchor = anchor_at(rho)
            history.append(current)
            times.append(now())
            if log and k % 20 == 0:
                log(f"surrogate iteration {k}: T = {current:.5f}, m {m}, fallbacks {fallbacks}")
        final = current
    torch.cuda.synchronize()
    counts = {key: evaluator.counts[key] - counts0.get(key, 0) for key in evaluator.counts}
    return {"final": final, "history": history, "times": times, "rho": rho,
            "fallbacks": fallbacks, "accepted": accepted, "inner_sizes": sizes,
            "errors": errors, "seconds": now(), "counts": counts}


def time_to(result, level):
    """Wall seconds until a loop's verified T first reached ``level`` (None if
    it never did)."""
    for t, value in zip(result["times"], result["history"]):
        if value >= level:
            return t
    return None
