"""Fourier-neural-operator surrogate of the density-design objective.

Maps a density rho (B, S, S) to the modal transmission at ``outputs``
wavelengths and to dT/drho (S, S) at the objective wavelength.  No
reference counterpart.  Every contraction runs in ``wfd_gemm_bf16`` on the
BF16 matrix cores with FP32 accumulation (weights kept as FP32 masters with a
BF16 copy for the products); elementwise, reduction and optimizer steps are
the library's own kernels; torch tensors are only storage.

Architecture (channels-last activations (B, S, S, W)):

* lift: [rho, x, y] -> W channels (pointwise, linear);
* ``layers`` Fourier layers v <- act(K v + W v + b): K keeps kx in
  {0..m-1, S-m..S-1} and ky in {0..m-1} of the 2-D DFT (dense DFT-matrix
  products along y, then x), mixes channels per mode with a real 2W x 2W
  block acting on (Re, Im) (R-linear), and transforms back with the
  Hermitian (irfft) weights; act = GELU except on the last layer;
* gradient head: pointwise W -> H (GELU) -> 1;
* transmission head: spatial mean -> H (GELU) -> ``outputs``.
"""

import ctypes

import numpy as np
import torch

from ._lib import call, lib, stream

P = ctypes.c_void_p


def _p(t):
    return P(t.data_ptr())


class Layout:
    """Offsets of every parameter in one flat FP32 buffer."""

    def __init__(self, width, modes, hidden, outputs, layers):
        W, m, H = width, modes, hidden
        self.entries = {}
        off = 0

        def add(name, shape):
            nonlocal off
            size = int(np.prod(shape))
            self.entries[name] = (off, tuple(shape))
            off += size

        add("lift_w", (3, W))
        add("lift_b", (W,))
        for l in range(layers):
            add(f"mix{l}", (2 * m * m, 2 * W, 2 * W))
            add(f"w{l}", (W, W))
            add(f"b{l}", (W,))
        add("g1_w", (W, H))
        add("g1_b", (H,))
        add("g2_w", (H, 1))
        add("g2_b", (1,))
        add("t1_w", (W, H))
        add("t1_b", (H,))
        add("t2_w", (H, outputs))
        add("t2_b", (outputs,))
        self.size = off

    def view(self, flat, name):
        off, shape = self.entries[name]
        return flat[off:off + int(np.prod(shape))].view(shape)


def dft_matrices(S, m, dtype=np.float32):
    """Forward/inverse DFT factors of the truncated transform (float64
    arithmetic, stored as ``dtype``).

    fy (2m, S): rows cos(2 pi ky y/S) then -sin (ky < m);
    fx (4m, 2S): (Re, Im) of sum_x exp(-i 2 pi kx x/S) over rows 2x + part,
        kx in {0..m-1, S-m..S-1};
    gx (2S, 4m): the inverse along x (exp(+i ...)), rows 2x + part;
    gy (S, 2m): Re(exp(+i 2 pi ky y/S) * w_ky) / S^2 with w_0 = 1, w_ky = 2.
    """
    if S < 2 * m:
        raise ValueError("the grid side must be at least twice the mode count")
    y = np.arange(S)
    ky = np.arange(m)
    kx = np.concatenate([np.arange(m), np.arange(S - m, S)])
    th = 2 * np.pi * np.outer(ky, y) / S                      # (m, S)
    fy = np.concatenate([np.cos(th), -np.sin(th)])           # (2m, S)
    ph = 2 * np.pi * np.outer(kx, y) / S                      # (2m, S)
    fx = np.zeros((4 * m, 2 * S))
    fx[:2 * m, 0::2] = np.cos(ph)
    fx[:2 * m, 1::2] = np.sin(ph)
    fx[2 * m:, 0::2] = -np.sin(ph)
    fx[2 * m:, 1::2] = np.cos(ph)
    gx = np.zeros((2 * S, 4 * m))
    gx[0::2, :2 * m] = np.cos(ph).T
    gx[0::2, 2 * m:] = -np.sin(ph).T
    gx[1::2, :2 * m] = np.sin(ph).T
    gx[1::2, 2 * m:] = np.cos(ph).T
    w = np.where(ky == 0, 1.0, 2.0)
    gy = np.concatenate([(w[:, None] * np.cos(th)).T, (-w[:, None] * np.sin(th)).T],
                        axis=1) / S ** 2                     # (S, 2m)
    return {k: np.ascontiguousarray(v, dtype) for k, v in
            (("fy", fy), ("fx", fx), ("gx", gx), ("gy", gy))}


def init_parameters(layout, width, modes, seed=0):
    """Seeded initial weights (uniform, fan-in scaled), biases zero."""
    rng = np.random.default_rng(seed)
    flat = np.zeros(layout.size, np.float32)
    for name, (off, shape) in layout.entries.items():
        if name.endswith("_b") or (name.startswith("b") and name[1:].isdigit()):
            continue
        fan_in = shape[-2] if len(shape) >= 2 else 1
        scale = 1.0 / np.sqrt(fan_in)
        if name.startswith("mix"):
            scale = 1.0 / (2 * width)
        flat[off:off + int(np.prod(shape))] = rng.uniform(-scale, scale, int(np.prod(shape)))
    return flat


def _splits(rows, most=32):
    """The largest divisor of ``rows`` up to ``most``: row runs per sample of
    a weight-gradient product."""
    return max(d for d in range(1, most + 1) if rows % d == 0)


class FNO:
    """The surrogate on libwfdtd kernels (training and inference)."""

    def __init__(self, side, width=32, modes=12, hidden=64, outputs=8, layers=4, batch=32,
                 device="cuda:0", seed=0):
        self.S, self.W, self.m, self.H, self.F, self.L = side, width, modes, hidden, outputs, layers
        self.B = batch
        self.device = torch.device(device)
        self.layout = Layout(width, modes, hidden, outputs, layers)
        dev = self.device
        self.params = torch.from_numpy(init_parameters(self.layout, width, modes, seed)).to(dev)
        self.params16 = torch.empty(self.layout.size, dtype=torch.bfloat16, device=dev)
        self.grads = torch.zeros(self.layout.size, device=dev)
        self.adam_m = torch.zeros(self.layout.size, device=dev)
        self.adam_v = torch.zeros(self.layout.size, device=dev)
        self.t = 0
        self.sync_bf16()
        mats = dft_matrices(side, modes)
        self.mats = {k: torch.from_numpy(v).to(dev) for k, v in mats.items()}
        S, W, m, B = side, width, modes, batch
        f32 = dict(device=dev, dtype=torch.float32)
        self.coords = torch.from_numpy(np.stack(np.meshgrid(
            np.linspace(0, 1, S), np.linspace(0, 1, S), indexing="ij"), -1).astype(np.float32)).to(dev)
        self.x_in = torch.zeros(B, S, S, 3, **f32)
        self.acts = [torch.zeros(B, S, S, W, **f32) for _ in range(layers + 1)]
        self.pres = [torch.zeros(B, S, S, W, **f32) for _ in range(layers)]
        self.y1 = torch.zeros(B * S, 2 * m, W, **f32)
        self.z = torch.zeros(B, 4 * m, m, W, **f32)
        self.o = torch.zeros(B, 4 * m, m, W, **f32)
        self.v1 = torch.zeros(B, 2 * S, m, W, **f32)
        self.s = torch.zeros(B, S, S, W, **f32)
        self.g_pre = torch.zeros(B, S, S, self.H, **f32)
        self.g_hid = torch.zeros(B, S, S, self.H, **f32)
        self.g_out = torch.zeros(B, S, S, 1, **f32)
        self.pool = torch.zeros(B, W, **f32)
        self.t_pre = torch.zeros(B, self.H, **f32)
        self.t_hid = torch.zeros(B, self.H, **f32)
        self.t_out = torch.zeros(B, outputs, **f32)
        # backward buffers
        self.d_act = torch.zeros(B, S, S, W, **f32)
        self.d_s = torch.zeros(B, S, S, W, **f32)
        self.d_v1 = torch.zeros(B, 2 * S, m, W, **f32)
        self.d_o = torch.zeros(B, 4 * m, m, W, **f32)
        self.d_z = torch.zeros(B, 4 * m, m, W, **f32)
        self.d_y1 = torch.zeros(B * S, 2 * m, W, **f32)
        self.d_ghid = torch.zeros(B, S, S, self.H, **f32)
        self.d_thid = torch.zeros(B, self.H, **f32)
        self.d_pool = torch.zeros(B, W, **f32)
        widest = max(W * W, self.H * max(W, outputs), 3 * W)
        self.partials = torch.zeros(B * _splits(S * S), widest, **f32)
        rows = B * S * S
        need = max(lib().wfd_colsum_scratch(rows, max(self.H, W)),
                   lib().wfd_colsum_scratch(B * _splits(S * S), widest),
                   lib().wfd_pool_mean_scratch(B, S * S, W))
        self.scratch = torch.zeros(int(need), **f32)

    # ------------------------------------------------------------ helpers
    def sync_bf16(self):
        call("wfd_to_bf16", _p(self.params), _p(self.params16), self.layout.size, stream())

    def p(self, name, bf16=False):
        return self.layout.view(self.params16 if bf16 else self.params, name)

    def gview(self, name):
        return self.layout.view(self.grads, name)

    def _w16(self, name):
        return self.layout.view(self.params16, name)

    @staticmethod
    def gemm(a, a16, sa, b, b16, sb, c, sc, batch, m, n, k, alpha=1.0, beta=0.0,
             a_off=0, b_off=0, c_off=0):
        ea = 2 if a16 else 4
        eb = 2 if b16 else 4
        call("wfd_gemm_bf16", P(a.data_ptr() + a_off * ea), int(a16), *sa,
             P(b.data_ptr() + b_off * eb), int(b16), *sb, P(c.data_ptr() + c_off * 4), *sc,
             int(batch), int(m), int(n), int(k), float(alpha), float(beta), stream())

    def linear(self, x, rows, k, w_name, n, out, beta=0.0, x_off=0):
        """out (rows, n) = x (rows, k) . W (k, n) [+ beta out]."""
        self.gemm(x, False, (0, k, 1), self._w16(w_name), True, (0, n, 1), out, (0, n, 1), 1,
                  rows, n, k, beta=beta, a_off=x_off)

    def colsum(self, x, rows, n, out, beta=1.0):
        call("wfd_colsum", _p(x), int(rows), int(n), _p(out), _p(self.scratch),
             self.scratch.numel(), float(beta), stream())

    # ------------------------------------------------------------ spectral
    def _spectral(self, x, out):
        """out (B,S,S,W) = irfft2(mix(rfft2(x) truncated))."""
        S, W, m, B = self.S, self.W, self.m, self.B
        M = self.mats
        # y-DFT: y1[b,x] (2m, W) = fy (2m, S) . x[b,x] (S, W)
        self.gemm(M["fy"], False, (0, S, 1), x, False, (S * W, W, 1), self.y1,
                  (2 * m * W, W, 1), B * S, 2 * m, W, S)
        # x-DFT: z[b] (4m, m*W) = fx (4m, 2S) . y1[b] (2S, m*W)
        self.gemm(M["fx"], False, (0, 2 * S, 1), self.y1, False, (2 * S * m * W, m * W, 1),
                  self.z, (4 * m * m * W, m * W, 1), B, 4 * m, m * W, 2 * S)
        return out

    def _mix(self, layer, src, dst, transpose=False):
        """dst parts = sum over src parts of src_part . block (per mode).
        src/dst (B, 4m, m, W): rows part*2m + kxi; mode k = kxi*m + ky."""
        m, W, B = self.m, self.W, self.B
        modes = 2 * m * m
        mix = self._w16(f"mix{layer}")
        part = 2 * m * m * W  # offset of the Im part
        for po in range(2):
            for pi in range(2):
                if not transpose:
                    # block rows pi*W.., cols po*W..
                    b_off = pi * W * 2 * W + po * W
                    sb = (4 * W * W, 2 * W, 1)
                    s_off, d_off = pi * part, po * part
                else:
                    # dst_pi += src_po . block(pi, po)^T
                    b_off = pi * W * 2 * W + po * W
                    sb = (4 * W * W, 1, 2 * W)
                    s_off, d_off = po * part, pi * part
                beta = 0.0 if ((pi == 0) if not transpose else (po == 0)) else 1.0
                self.gemm(src, False, (W, 4 * m * m * W, 1), mix, True, sb, dst,
                          (W, 4 * m * m * W, 1), modes, B, W, W, beta=beta, a_off=s_off,
                          b_off=b_off, c_off=d_off)

    def _inverse(self, o, out):
        S, W, m, B = self.S, self.W, self.m, self.B
        M = self.mats
        # x-inverse: v1[b] (2S, m*W) = gx (2S, 4m) . o[b] (4m, m*W)
        self.gemm(M["gx"], False, (0, 4 * m, 1), o, False, (4 * m * m * W, m * W, 1), self.v1,
                  (2 * S * m * W, m * W, 1), B, 2 * S, m * W, 4 * m)
        # y-inverse: out[b,x] (S, W) = gy (S, 2m) . v1[b,x] (2m, W)
        self.gemm(M["gy"], False, (0, 2 * m, 1), self.v1, False, (2 * m * W, W, 1), out,
                  (S * W, W, 1), B * S, S, W, 2 * m)

    # ------------------------------------------------------------ forward
    def forward(self, rho):
        """rho (B, S, S) float32 device tensor -> (T (B, outputs), g (B, S, S))."""
        S, W, B, H = self.S, self.W, self.B, self.H
        if tuple(rho.shape) != (B, S, S):
            raise ValueError(f"rho must have shape {(B, S, S)}")
        rows = B * S * S
        self.x_in[..., 0].copy_(rho)
        self.x_in[..., 1:].copy_(self.coords.expand(B, S, S, 2))
        self.linear(self.x_in, rows, 3, "lift_w", W, self.acts[0])
        call("wfd_bias_act", _p(self.acts[0]), _p(self.p("lift_b")), None, _p(self.acts[0]),
             rows, W, 0, stream())
        for l in range(self.L):
            x = self.acts[l]
            self._spectral(x, self.s)
            self._mix(l, self.z, self.o)
            self._inverse(self.o, self.s)
            self.linear(x, rows, W, f"w{l}", W, self.s, beta=1.0)
            act = 1 if l + 1 < self.L else 0
            call("wfd_bias_act", _p(self.s), _p(self.p(f"b{l}")), _p(self.pres[l]),
                 _p(self.acts[l + 1]), rows, W, act, stream())
        top = self.acts[self.L]
        # gradient head
        self.linear(top, rows, W, "g1_w", H, self.g_pre)
        call("wfd_bias_act", _p(self.g_pre), _p(self.p("g1_b")), _p(self.g_pre), _p(self.g_hid),
             rows, H, 1, stream())
        self.linear(self.g_hid, rows, H, "g2_w", 1, self.g_out)
        call("wfd_bias_act", _p(self.g_out), _p(self.p("g2_b")), None, _p(self.g_out), rows, 1,
             0, stream())
        # transmission head
        call("wfd_pool_mean", _p(top), B, S * S, W, _p(self.pool), _p(self.scratch),
             self.scratch.numel(), stream())
        self.linear(self.pool, B, W, "t1_w", H, self.t_pre)
        call("wfd_bias_act", _p(self.t_pre), _p(self.p("t1_b")), _p(self.t_pre),
             _p(self.t_hid), B, H, 1, stream())
        self.linear(self.t_hid, B, H, "t2_w", self.F, self.t_out)
        call("wfd_bias_act", _p(self.t_out), _p(self.p("t2_b")), None, _p(self.t_out), B,
             self.F, 0, stream())
        return self.t_out, self.g_out.view(B, S, S)

    # ------------------------------------------------------------ backward
    def _wgrad(self, x, rows_per, k, dy, n, name):
        """grad[name] (k, n) = sum_b x_b^T (k x rows) . dy_b (rows x n).  Each
        sample's rows are split into ``splits`` equal runs (one GEMM batch
        entry and one partial product each, so the small k x n products of a
        10,000-row sample spread over hundreds of workgroups instead of one
        per sample), and the partials are summed in a fixed order."""
        B = self.B
        splits = _splits(rows_per)
        R = rows_per // splits
        part = self.partials.view(-1)[:B * splits * k * n].view(B * splits, k * n)
        self.gemm(x, False, (R * k, 1, k), dy, False, (R * n, n, 1), part,
                  (k * n, n, 1), B * splits, k, n, R)
        self.colsum(part, B * splits, k * n, self.gview(name).view(-1), beta=0.0)

    def backward(self, d_t, d_g):
        """Accumulate parameter gradients of sum(d_t * T) + sum(d_g * g)
        (overwrites self.grads)."""
        S, W, B, H, m = self.S, self.W, self.B, self.H, self.m
        rows = B * S * S
        s = stream()
        top = self.acts[self.L]
        # transmission head
        d_t = d_t.contiguous()
        self.colsum(d_t, B, self.F, self.gview("t2_b"), beta=0.0)
        self._wgrad(self.t_hid, 1, H, d_t, self.F, "t2_w")
        self.gemm(d_t, False, (0, self.F, 1), self._w16("t2_w"), True, (0, 1, self.F),
                  self.d_thid, (0, H, 1), 1, B, H, self.F)
        call("wfd_act_back", _p(self.d_thid), _p(self.t_pre), _p(self.d_thid), B * H, 1, s)
        self.colsum(self.d_thid, B, H, self.gview("t1_b"), beta=0.0)
        self._wgrad(self.pool, 1, W, self.d_thid, H, "t1_w")
        self.gemm(self.d_thid, False, (0, H, 1), self._w16("t1_w"), True, (0, 1, H),
                  self.d_pool, (0, W, 1), 1, B, W, H)
        call("wfd_pool_mean_back", _p(self.d_pool), B, S * S, W, _p(self.d_act), 0.0, s)
        # gradient head
        d_g = d_g.contiguous()
        self.colsum(d_g, rows, 1, self.gview("g2_b"), beta=0.0)
        self._wgrad(self.g_hid, S * S, H, d_g, 1, "g2_w")
        self.gemm(d_g, False, (0, 1, 1), self._w16("g2_w"), True, (0, 1, 1), self.d_ghid,
                  (0, H, 1), 1, rows, H, 1)
        call("wfd_act_back", _p(self.d_ghid), _p(self.g_pre), _p(self.d_ghid), rows * H, 1, s)
        self.colsum(self.d_ghid, rows, H, self.gview("g1_b"), beta=0.0)
        self._wgrad(top, S * S, W, self.d_ghid, H, "g1_w")
        self.gemm(self.d_ghid, False, (0, H, 1), self._w16("g1_w"), True, (0, 1, H), self.d_act,
                  (0, W, 1), 1, rows, W, H, beta=1.0)
        # Fourier layers, last to first
        for l in reversed(range(self.L)):
            act = 1 if l + 1 < self.L else 0
            call("wfd_act_back", _p(self.d_act), _p(self.pres[l]), _p(self.d_s), rows * W, act, s)
            self.colsum(self.d_s, rows, W, self.gview(f"b{l}"), beta=0.0)
            x = self.acts[l]
            self._wgrad(x, S * S, W, self.d_s, W, f"w{l}")
            # d x = d_s . W^T  (pointwise path)
            self.gemm(self.d_s, False, (0, W, 1), self._w16(f"w{l}"), True, (0, 1, W),
                      self.d_act, (0, W, 1), 1, rows, W, W)
            # spectral path: transposes of y-inverse, x-inverse
            M = self.mats
            self.gemm(M["gy"], False, (0, 1, 2 * m), self.d_s, False, (S * W, W, 1), self.d_v1,
                      (2 * m * W, W, 1), B * S, 2 * m, W, S)
            self.gemm(M["gx"], False, (0, 1, 4 * m), self.d_v1, False, (2 * S * m * W, m * W, 1),
                      self.d_o, (4 * m * m * W, m * W, 1), B, 4 * m, m * W, 2 * S)
            # mix weights: per mode, block(pi, po) grad = z_pi^T . d_o_po
            self._mix_wgrad(l)
            self._mix(l, self.d_o, self.d_z, transpose=True)
            self.gemm(M["fx"], False, (0, 1, 2 * S), self.d_z, False, (4 * m * m * W, m * W, 1),
                      self.d_y1, (2 * S * m * W, m * W, 1), B, 2 * S, m * W, 4 * m)
            self.gemm(M["fy"], False, (0, 1, S), self.d_y1, False, (2 * m * W, W, 1), self.d_act,
                      (S * W, W, 1), B * S, S, W, 2 * m, beta=1.0)
            # the spectral forward of layer l recomputed z from acts[l]; the
            # stored z is the last layer's, so recompute it for the weight grads
        # lift
        self.colsum(self.d_act, rows, W, self.gview("lift_b"), beta=0.0)
        self._wgrad(self.x_in, S * S, 3, self.d_act, W, "lift_w")

    def _mix_wgrad(self, layer):
        m, W, B = self.m, self.W, self.B
        modes = 2 * m * m
        # z of this layer: recompute from acts[layer]
        self._spectral(self.acts[layer], None)
        g = self.gview(f"mix{layer}")
        part = 2 * m * m * W
        for pi in range(2):
            for po in range(2):
                # grad block (pi, po) [mode] (W x W) = z_pi[mode]^T (W x B) . d_o_po[mode] (B x W)
                self.gemm(self.z, False, (W, 1, 4 * m * m * W), self.d_o, False,
                          (W, 4 * m * m * W, 1), g, (4 * W * W, 2 * W, 1), modes, W, W, B,
                          a_off=pi * part, b_off=po * part, c_off=pi * W * 2 * W + po * W)

    # ------------------------------------------------------------ training
    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8, decay=0.0):
        self.t += 1
        call("wfd_adam", _p(self.params), _p(self.grads), _p(self.adam_m), _p(self.adam_v),
             _p(self.params16), self.layout.size, float(lr), float(b1), float(b2), float(eps),
             self.t, float(decay), stream())
