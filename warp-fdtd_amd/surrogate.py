"""Surrogate-assisted inverse design of the density bend (no reference counterpart).

* ``surrogate_case`` writes the bend bundle the surrogate learns: modal DFT
  planes at 8 wavelengths (1.50-1.60 um), a side x side design region.
* ``Evaluator`` runs the solver: batched forward + adjoint labels
  (``BatchTransmission``: T spectrum and dT/drho at the objective
  wavelength), single forward verifications (a plain ``Solver``, resident
  when that is faster) and single adjoint steps.
* ``train`` fits the FNO (fno.py) to (rho -> T spectrum, dT/drho) on the
  library's own BF16 kernels.
* ``optimize`` runs gradient ascent on T at the objective wavelength either
  with the adjoint at every step (``surrogate=None``) or with surrogate
  proposals, each re-verified by a full FDTD forward and replaced by an
  adjoint step when the surrogate's transmission error exceeds ``tolerance``
  (the surrogate's T is corrected by the last verified design's: only the
  change from it is predicted).
"""

import time

import numpy as np
import torch

from . import detectors
from .cases import bend_case, design_density
from .design import BatchTransmission, DensityMap, _design_box
from .solver import Solver

WAVELENGTHS = np.linspace(1.50e-6, 1.60e-6, 8)


def surrogate_case(path, side=100, spacing=40e-9, steps=1500, pml=8, margin=30, nz=24,
                   pulse_duration=30e-15):
    """Bend with a side x side voxel design region (spacing per voxel), the
    arms and planes ``margin`` cells outside it, modal planes at 8 wavelengths."""
    n = side + 2 * (pml + margin)
    return bend_case(path, np.full((side, side), 0.5), shape=(n, n, nz), spacing=spacing,
                     steps=steps, pml=pml, design=(side, side), wavelengths=list(WAVELENGTHS),
                     wavelength=1.55e-6, pulse_duration=pulse_duration)


def objective_index():
    """The sampled wavelength nearest 1.55 um (the objective's)."""
    return int(np.argmin(np.abs(WAVELENGTHS - 1.55e-6)))


class Evaluator:
    """FDTD evaluations of densities on one surrogate bundle."""

    def __init__(self, case, batch=32, checkpoint_every=192, device="cuda:0", log=None):
        self.log = log or (lambda message: None)
        self.case = case
        self.wavelength = float(WAVELENGTHS[objective_index()])
        self.batch = batch
        self.labeler = BatchTransmission([case] * batch, wavelength=self.wavelength,
                                         checkpoint_every=checkpoint_every, device=device)
        self.labeler.log = self.log
        self.log(f"compiling {batch} batched adjoints")
        self.labeler.compile()
        self.single = self.labeler.problems[0]
        self.checker = Solver(case, device=device)
        detectors.attach(self.checker)
        self.checker.compile()
        self.map = DensityMap(_design_box(case))
        self.side = self.map.plane[0]
        self.counts = {"forward": 0, "adjoint": 0, "labels": 0}

    def label(self, rhos):
        """(T spectra (n, 8), dT/drho (n, side, side)) for the densities."""
        rhos = list(rhos)
        spectra, grads = [], []
        for start in range(0, len(rhos), self.batch):
            chunk = rhos[start:start + self.batch]
            pad = chunk + [chunk[-1]] * (self.batch - len(chunk))
            self.labeler.forward(pad)
            g = self.labeler.backward()
            self.log(f"labelled {min(start + self.batch, len(rhos))}/{len(rhos)} designs")
            for k in range(len(chunk)):
                spectra.append(self.labeler.problems[k].adjoint.transmission_spectrum())
                grads.append(g[k].cpu().numpy())
        self.counts["labels"] += len(rhos)
        return np.array(spectra), np.array(grads)

    def transmission(self, rho):
        """T spectrum by one full FDTD forward (the verification)."""
        s = self.checker
        self.map.write(rho, s)
        s.reset()
        s.run()
        self.counts["forward"] += 1
        return detectors.dft_transmission(s)

    def adjoint(self, rho):
        """(T spectrum, dT/drho at the objective wavelength) by forward + adjoint."""
        T, g = self.single.value_and_grad(rho)
        self.counts["adjoint"] += 1
        return self.single.adjoint.transmission_spectrum(), g.cpu().numpy()


def random_designs(count, side, seed=0):
    """Seeded blurred-random densities (sigma 2 voxels)."""
    rng = np.random.default_rng(seed)
    return [design_density(rng, (side, side), blurred=True) for _ in range(count)]


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


def optimize(evaluator, rho0, steps=200, step=0.05, surrogate=None, tolerance=0.01, log=None,
             predictor="taylor", require_improvement=True, tuner=None, anchored=True,
             lagged=False):
    """Gradient ascent on T at the objective wavelength.  ``surrogate`` =
    (net, norm) proposes each step; every proposal is re-verified by a full
    FDTD forward, and a proposal whose surrogate T is off by more than
    ``tolerance`` (relative) is replaced by an adjoint step.

    Gradients of a proposal step: with ``anchored`` the exact adjoint
    gradient of the last adjoint design (the anchor) plus the network's
    change of gradient since it, g_a + g_net(rho) - g_net(rho_a) (the first
    step takes an adjoint step to get an anchor); else the network's g_net.
    ``lagged`` drops the network's correction (g_a alone: the ablation
    without a surrogate).  The surrogate's T of a proposal (``predictor``),
    with T0 the last verified T and d = proposal - current design:
    "taylor": T0 + <g(current) + g(proposal), d> / 2 (trapezoid rule);
    "difference": T0 + net(proposal) - net(current) with the network's T
    head; "absolute": net(proposal).  ``require_improvement``: a proposal
    whose verified T is below the current one is replaced by an adjoint step
    as well, so accepted surrogate steps never lose ground.  ``tuner`` (an
    OnlineTuner) learns from every adjoint label the loop computes.
    Returns a dict with the T trajectory (verified), the surrogate's relative
    T error on every proposal, counts and wall time."""
    f = objective_index()
    rho = np.array(rho0, np.float64)
    start = time.perf_counter()
    history, errors, fallbacks = [], [], 0
    if surrogate is None:
        for k in range(steps):
            spec, g = evaluator.adjoint(rho)
            history.append(float(spec[f]))
            rho = _ascent(rho, g, step)
            if log and k % 20 == 0:
                log(f"adjoint step {k}: T = {history[-1]:.5f}")
        final = float(evaluator.transmission(rho)[f])
    else:
        net, norm = surrogate
        anchor = None  # (network gradient at the anchor, adjoint gradient there)

        def gradient(pred):
            if not anchored:
                return pred[1]
            if lagged:
                return anchor[1]
            return anchor[1] + pred[1] - anchor[0]

        def adjoint_step(rho):
            nonlocal anchor
            spec, g = evaluator.adjoint(rho)
            if tuner is not None:
                tuner.add(rho, spec, g)
            anchor = (predict(net, norm, rho)[1], g)
            return _ascent(rho, g, step)

        current = float(evaluator.transmission(rho)[f])
        here = predict(net, norm, rho)
        for _ in range(steps):
            if anchored and anchor is None:
                fallbacks += 1
                proposal = adjoint_step(rho)
                t_true = float(evaluator.transmission(proposal)[f])
                there = predict(net, norm, proposal)
            else:
                g_here = gradient(here)
                proposal = _ascent(rho, g_here, step)
                there = predict(net, norm, proposal)
                d = proposal - rho
                if predictor == "taylor":
                    t_hat = current + 0.5 * float(((g_here + gradient(there)) * d).sum())
                elif predictor == "difference":
                    t_hat = current + there[0][f] - here[0][f]
                else:
                    t_hat = there[0][f]
                t_true = float(evaluator.transmission(proposal)[f])
                errors.append(abs(t_hat - t_true) / max(abs(t_true), 1e-12))
                if errors[-1] > tolerance or (require_improvement and t_true < current):
                    fallbacks += 1
                    proposal = adjoint_step(rho)
                    t_true = float(evaluator.transmission(proposal)[f])
                    there = predict(net, norm, proposal)
            rho, current, here = proposal, t_true, there
            history.append(current)
            if log and len(history) % 20 == 1:
                log(f"surrogate step {len(history) - 1}: T = {current:.5f}, fallbacks {fallbacks}")
        final = current
    torch.cuda.synchronize()
    return {"final": final, "history": history, "rho": rho, "fallbacks": fallbacks,
            "errors": errors, "seconds": time.perf_counter() - start,
            "counts": dict(evaluator.counts)}
