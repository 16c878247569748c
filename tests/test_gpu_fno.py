"""BF16 matrix-core kernels of the FNO surrogate against float64 references."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _bf16(x):
    """Round to BF16 (nearest even) and back, on the CPU in float32."""
    return torch.as_tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float32)


def _gemm(pkg, A, B, C, alpha=1.0, beta=0.0, a_bf16=False, b_bf16=False, sa=None, sb=None):
    """Run wfd_gemm_bf16 on batched (b, m, k) x (b, k, n) with the given stride
    tuples (elements) or the contiguous ones."""
    from ctypes import c_void_p
    lib = pkg._lib
    bat, m, k = sa[0] if sa else A.shape
    n = C.shape[2]
    sa = sa[1] if sa else (m * k, k, 1)
    sb = sb[1] if sb else (k * n, n, 1)
    lib.call("wfd_gemm_bf16", c_void_p(A.data_ptr()), int(a_bf16), *sa, c_void_p(B.data_ptr()),
             int(b_bf16), *sb, c_void_p(C.data_ptr()), m * n, n, 1, bat, m, n, k,
             float(alpha), float(beta), lib.stream())
    torch.cuda.synchronize()


@pytest.mark.parametrize("shape", [(1, 64, 64, 32), (3, 100, 24, 37), (2, 7, 130, 300),
                                   (5, 129, 65, 64)])
def test_gemm_matches_float64_of_bf16_operands(pkg, gpu, shape):
    """Ragged tiles, odd K: FP32 accumulation of BF16-rounded operands agrees
    with a float64 product of the same rounded operands to 1e-5 (relative)."""
    bat, m, n, k = shape
    g = torch.Generator().manual_seed(sum(shape))
    A = torch.randn(bat, m, k, generator=g)
    B = torch.randn(bat, k, n, generator=g)
    C0 = torch.randn(bat, m, n, generator=g)
    want = 0.5 * (_bf16(A).double() @ _bf16(B).double()) + 2.0 * C0.double()
    C = C0.cuda()
    _gemm(pkg, A.cuda(), B.cuda(), C, alpha=0.5, beta=2.0)
    err = (C.cpu().double() - want).norm() / want.norm()
    assert err < 1e-5, err


def test_gemm_strides_transposes_and_bf16_storage(pkg, gpu):
    """A^T and B^T through stride swaps, BF16-stored operands: same result."""
    g = torch.Generator().manual_seed(1)
    bat, m, n, k = 2, 45, 70, 96
    At = torch.randn(bat, k, m, generator=g)   # stored (k, m): A = At^T
    Bt = torch.randn(bat, n, k, generator=g)   # stored (n, k): B = Bt^T
    want = _bf16(At).double().transpose(1, 2) @ _bf16(Bt).double().transpose(1, 2)
    C = torch.zeros(bat, m, n).cuda()
    _gemm(pkg, At.cuda(), Bt.cuda(), C, sa=((bat, m, k), (k * m, 1, m)),
          sb=(None, (n * k, 1, k)))
    assert (C.cpu().double() - want).norm() / want.norm() < 1e-5
    # BF16 bit patterns in memory give the same bits as FP32 operands
    A16 = At.transpose(1, 2).contiguous().to(torch.bfloat16).cuda()
    B16 = Bt.transpose(1, 2).contiguous().to(torch.bfloat16).cuda()
    C16 = torch.zeros(bat, m, n).cuda()
    _gemm(pkg, A16, B16, C16, a_bf16=True, b_bf16=True)
    np.testing.assert_array_equal(C16.cpu().numpy(), C.cpu().numpy())


def test_elementwise_and_reductions(pkg, gpu):
    """bias+GELU and its derivative, column sums, mean pooling, MSE, Adam."""
    from ctypes import c_void_p as P
    lib = pkg._lib
    s = lib.stream()
    g = torch.Generator().manual_seed(2)
    rows, n = 3000, 37
    y = torch.randn(rows, n, generator=g)
    bias = torch.randn(n, generator=g)
    pre, out = torch.empty(rows, n).cuda(), torch.empty(rows, n).cuda()
    yd, bd = y.cuda(), bias.cuda()
    lib.call("wfd_bias_act", P(yd.data_ptr()), P(bd.data_ptr()), P(pre.data_ptr()),
             P(out.data_ptr()), rows, n, 1, s)
    z = (y + bias).double()
    gelu = 0.5 * z * (1 + torch.erf(z / np.sqrt(2)))
    assert (out.cpu().double() - gelu).abs().max() < 1e-5
    dout = torch.randn(rows, n, generator=g)
    dz = torch.empty(rows, n).cuda()
    dout_d = dout.cuda()
    lib.call("wfd_act_back", P(dout_d.data_ptr()), P(pre.data_ptr()), P(dz.data_ptr()),
             rows * n, 1, s)
    zz = z.clone().requires_grad_()
    (0.5 * zz * (1 + torch.erf(zz / np.sqrt(2)))).backward(dout.double())
    torch.cuda.synchronize()
    assert (dz.cpu().double() - zz.grad).abs().max() < 1e-5
    scratch = torch.empty(lib.lib().wfd_colsum_scratch(rows, n)).cuda()
    cs = torch.ones(n).cuda()
    lib.call("wfd_colsum", P(yd.data_ptr()), rows, n, P(cs.data_ptr()), P(scratch.data_ptr()),
             scratch.numel(), 1.0, s)
    torch.cuda.synchronize()
    assert (cs.cpu().double() - (y.double().sum(0) + 1)).abs().max() < 1e-3
    x = torch.randn(4, 250, 9, generator=g)
    pooled = torch.empty(4, 9).cuda()
    x_d = x.cuda()
    pscratch = torch.empty(lib.lib().wfd_pool_mean_scratch(4, 1250, 9)).cuda()
    x = torch.randn(4, 1250, 9, generator=g)  # (three runs of pixels and a ragged one)
    x_d = x.cuda()
    lib.call("wfd_pool_mean", P(x_d.data_ptr()), 4, 1250, 9, P(pooled.data_ptr()),
             P(pscratch.data_ptr()), pscratch.numel(), s)
    torch.cuda.synchronize()
    assert (pooled.cpu().double() - x.double().mean(1)).abs().max() < 1e-6
    pred, target = torch.randn(1000, generator=g), torch.randn(1000, generator=g)
    d = torch.empty(1000).cuda()
    parts = torch.empty(lib.lib().wfd_mse_partials(1000)).cuda()
    pred_d, target_d = pred.cuda(), target.cuda()
    lib.call("wfd_mse", P(pred_d.data_ptr()), P(target_d.data_ptr()), 1000, 0.002,
             P(d.data_ptr()), P(parts.data_ptr()), s)
    torch.cuda.synchronize()
    assert abs(parts.cpu().double().sum() - ((pred - target).double() ** 2).sum()) < 1e-3
    assert (d.cpu() - 0.002 * (pred - target)).abs().max() < 1e-7
    w = torch.randn(500, generator=g)
    gr = torch.randn(500, generator=g)
    wd, m, v = w.cuda(), torch.zeros(500).cuda(), torch.zeros(500).cuda()
    w16 = torch.empty(500, dtype=torch.bfloat16).cuda()
    gr_d = gr.cuda()
    lib.call("wfd_adam", P(wd.data_ptr()), P(gr_d.data_ptr()), P(m.data_ptr()),
             P(v.data_ptr()), P(w16.data_ptr()), 500, 1e-2, 0.9, 0.999, 1e-8, 1, 0.0, s)
    ref = w.clone().requires_grad_()
    opt = torch.optim.Adam([ref], lr=1e-2)
    ref.grad = gr.clone()
    opt.step()
    torch.cuda.synchronize()
    assert (wd.cpu() - ref.detach()).abs().max() < 1e-6
    np.testing.assert_array_equal(w16.cpu().float().numpy(), wd.cpu().to(torch.bfloat16).float().numpy())


@pytest.mark.parametrize("S,B,outputs", [(32, 3, 8), (100, 2, 9)])
def test_fno_forward_and_backward_match_float64_oracle(pkg, gpu, S, B, outputs):
    """The GPU network (BF16 products, FP32 accumulation, hand-written
    backward) against the float64 CPU oracle (torch FFTs + autograd) with the
    same BF16-rounded weights: outputs and every parameter gradient, on a
    small grid and on the surrogate's 100x100 design grid with its 9
    transmission outputs."""
    from oracle import fno as OF
    net = pkg.fno.FNO(S, batch=B, outputs=outputs, seed=3)
    g = torch.Generator().manual_seed(4)
    rho = torch.rand(B, S, S, generator=g)
    d_t = torch.randn(B, net.F, generator=g)
    d_g = torch.randn(B, S, S, generator=g)
    T, gr = net.forward(rho.cuda())
    net.backward(d_t.cuda(), d_g.cuda())
    torch.cuda.synchronize()
    params = OF.unpack(net.params.cpu().numpy(), net.layout)
    T0, g0, grads = OF.gradients(params, rho.double(), d_t.double(), d_g.double(), net.L, net.m)
    rel = lambda a, b: float((a.cpu().double() - b).norm() / b.norm())
    assert rel(T, T0) < 2e-2, rel(T, T0)
    assert rel(gr, g0) < 2e-2, rel(gr, g0)
    worst = {}
    for name in net.layout.entries:
        got = net.gview(name).cpu().double()
        worst[name] = rel(got, grads[name].reshape(got.shape))
    assert max(worst.values()) < 5e-2, worst
