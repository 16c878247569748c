"""CPU float64 restatement of the FNO surrogate (TEST INFRASTRUCTURE ONLY).

No reference counterpart.  The same network as warp-fdtd_amd/fno.py, written
with torch's CPU FFTs (rfft2 / irfft2) instead of DFT-matrix products and with
torch.autograd for the backward pass, so it checks both the GPU forward and
the GPU's hand-written backward.  Parameters come in the product's flat
layout (``Layout``); ``bf16=True`` rounds the weights to BF16 first, as the
GPU's products see them.
"""

import numpy as np
import torch


def _round_bf16(t):
    return t.to(torch.float32).to(torch.bfloat16).to(torch.float64)


def unpack(flat, layout, bf16=True):
    flat = torch.as_tensor(np.asarray(flat), dtype=torch.float64)
    out = {}
    for name, (off, shape) in layout.entries.items():
        v = flat[off:off + int(np.prod(shape))].reshape(shape)
        # the GEMMs read weights as BF16; biases are added in FP32
        is_bias = name.endswith("_b") or (name.startswith("b") and name[1:].isdigit())
        out[name] = (_round_bf16(v) if (bf16 and not is_bias) else v).clone().requires_grad_()
    return out


def spectral(x, mix, m):
    """x (B, S, S, W) -> irfft2(per-mode R-linear mix of rfft2(x) truncated)."""
    B, S, _, W = x.shape
    X = torch.fft.rfft2(x.permute(0, 3, 1, 2), s=(S, S))          # (B, W, S, S//2+1)
    kx = torch.cat([torch.arange(m), torch.arange(S - m, S)])
    Z = X[:, :, kx, :m]                                            # (B, W, 2m, m)
    Z = Z.permute(0, 2, 3, 1).reshape(B, 2 * m * m, W)             # mode k = kxi*m + ky
    zc = torch.cat([Z.real, Z.imag], dim=-1)                       # (B, modes, 2W)
    oc = torch.einsum("bki,kio->bko", zc, mix)                     # (B, modes, 2W)
    O = torch.complex(oc[..., :W], oc[..., W:]).reshape(B, 2 * m, m, W).permute(0, 3, 1, 2)
    full = torch.zeros(B, W, S, S // 2 + 1, dtype=O.dtype)
    full[:, :, kx, :m] = O
    out = torch.fft.irfft2(full, s=(S, S))                         # (B, W, S, S)
    return out.permute(0, 2, 3, 1)


def forward(params, rho, layers, m):
    """rho (B, S, S) float64 -> (T (B, F), g (B, S, S))."""
    B, S, _ = rho.shape
    g = np.linspace(0, 1, S)
    xs, ys = np.meshgrid(g, g, indexing="ij")
    coords = torch.as_tensor(np.stack([xs, ys], -1), dtype=torch.float64)
    x = torch.cat([rho[..., None], coords.expand(B, S, S, 2)], dim=-1)
    v = x @ params["lift_w"] + params["lift_b"]
    gelu = torch.nn.functional.gelu
    for l in range(layers):
        z = spectral(v, params[f"mix{l}"], m) + v @ params[f"w{l}"] + params[f"b{l}"]
        v = gelu(z) if l + 1 < layers else z
    gh = gelu(v @ params["g1_w"] + params["g1_b"])
    grad = (gh @ params["g2_w"] + params["g2_b"])[..., 0]
    pooled = v.mean(dim=(1, 2))
    th = gelu(pooled @ params["t1_w"] + params["t1_b"])
    return th @ params["t2_w"] + params["t2_b"], grad


def gradients(params, rho, d_t, d_g, layers, m):
    """d/d(params) of sum(d_t * T) + sum(d_g * g)."""
    T, g = forward(params, rho, layers, m)
    total = (T * d_t).sum() + (g * d_g).sum()
    names = list(params)
    grads = torch.autograd.grad(total, [params[k] for k in names])
    return T.detach(), g.detach(), dict(zip(names, grads))
