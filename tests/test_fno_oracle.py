"""The FNO's truncated-DFT factorisation against the CPU oracle's FFTs (no GPU)."""

import numpy as np
import torch

from oracle import fno as OF


def _spectral_by_matrices(pkg, x, mix, m):
    """The GPU path's product order in float64: y-DFT, x-DFT, per-mode
    block mix, x-inverse, y-inverse (warp-fdtd_amd/fno.py)."""
    B, S, _, W = x.shape
    M = pkg.fno.dft_matrices(S, m, np.float64)
    y1 = np.einsum("ry,bxyc->bxrc", M["fy"], x)                        # (B, S, 2m, W)
    y1 = y1.reshape(B, S, 2, m, W).transpose(0, 1, 2, 3, 4).reshape(B, 2 * S, m * W)
    z = np.einsum("qk,bkn->bqn", M["fx"], y1).reshape(B, 2, 2 * m, m, W)  # part, kxi, ky, c
    zc = np.concatenate([z[:, 0], z[:, 1]], axis=-1).reshape(B, 2 * m * m, 2 * W)
    oc = np.einsum("bki,kio->bko", zc, mix)
    o = np.stack([oc[..., :W], oc[..., W:]], axis=1).reshape(B, 4 * m, m * W)
    v1 = np.einsum("kq,bqn->bkn", M["gx"], o).reshape(B, S, 2, m, W).reshape(B, S, 2 * m, W)
    return np.einsum("yr,bxrc->bxyc", M["gy"], v1)


def test_dft_matrix_factorisation_equals_rfft2(pkg):
    """rfft2 -> truncate -> R-linear mix -> irfft2 (oracle) equals the DFT-matrix
    products of the GPU path, for odd and even sides."""
    rng = np.random.default_rng(0)
    for S, m, W in ((24, 12, 3), (29, 6, 2), (40, 12, 4)):
        x = rng.normal(size=(2, S, S, W))
        mix = rng.normal(size=(2 * m * m, 2 * W, 2 * W))
        want = OF.spectral(torch.from_numpy(x), torch.from_numpy(mix), m).numpy()
        got = _spectral_by_matrices(pkg, x, mix, m)
        assert np.abs(got - want).max() < 1e-10 * max(1.0, np.abs(want).max())


def test_oracle_catches_a_dropped_conjugate_term(pkg):
    """A planted mistake (the imaginary rows of the x-DFT negated) is caught."""
    rng = np.random.default_rng(1)
    S, m, W = 24, 12, 2
    x = rng.normal(size=(1, S, S, W))
    mix = rng.normal(size=(2 * m * m, 2 * W, 2 * W))
    want = OF.spectral(torch.from_numpy(x), torch.from_numpy(mix), m).numpy()
    real = pkg.fno.dft_matrices

    def broken(S_, m_, dtype=np.float32):
        M = real(S_, m_, dtype)
        M["fx"] = M["fx"].copy()
        M["fx"][2 * m_:] *= -1
        return M

    pkg.fno.dft_matrices = broken
    try:
        got = _spectral_by_matrices(pkg, x, mix, m)
    finally:
        pkg.fno.dft_matrices = real
    assert np.abs(got - want).max() > 1e-3
