"""Minimum-variance beamforming written with plain PyTorch operations on a
whole frame at once (gather of the subarrays, batched outer products,
temporal sum, torch.linalg.solve): the plain per-frame GPU form that the
fused kernel is compared with, for accuracy in the GPU tests and for speed in
bench.py.

TEST AND BENCHMARK INFRASTRUCTURE: never imported by the supra_amd package.

Same per-sample algorithm as the reference's RxBeamformerMV
(RxBeamformerMV.cu:44-113 aperture and subarrays, :151-228 covariance and
temporal sum, :239-305 diagonal loading, :558-586 the solve, :307-378
weights), over the full depth: per sample the non-zero channel range
[f, l], L_t = min(L, l-f+1), K_act = min(K, l-f-L_t+2), R(t) = sum of the
K_act outer products, R~(t) = sum of R(tau) over |tau - t| <= T (each R(tau)
on its own aperture), delta = alpha trace/L_t on the L_t block, R~ w = 1,
y = N_ch sum_i (sum_k x_k,i) w_i / (sum w K_act^p), 0 if |y| > 1e7 or NaN.
"""
from __future__ import annotations

import torch


def plain_mv(x, sub_array_size: int, temporal_smoothing: int, scaling_power: float, loading_factor: float):
    """x [scanline][channel][sample] float32 -> image [sample][scanline] float32."""
    n_rx, n_ch, n_d = x.shape
    L = sub_array_size or n_ch // 2
    K = n_ch - L + 1
    T = temporal_smoothing
    dev = x.device
    xt = x.permute(0, 2, 1).contiguous()                          # [rx, d, ch]
    nz = xt != 0
    ar = torch.arange(n_ch, device=dev)
    f = torch.where(nz, ar, n_ch).amin(-1)
    f = torch.where(f == n_ch, 0, f)
    l = torch.where(nz, ar, -1).amax(-1).clamp(min=0)
    lt = torch.clamp(l - f + 1, max=L)
    kact = torch.clamp(torch.clamp(l - f - lt + 2, max=K), min=0)
    k = torch.arange(K, device=dev)
    i = torch.arange(L, device=dev)
    idx = (f[..., None, None] + k[:, None] + i[None, :]).clamp(max=n_ch - 1)     # [rx, d, K, L]
    X = torch.gather(xt[:, :, None, :].expand(n_rx, n_d, K, n_ch), 3, idx)
    X = X * ((k[:, None] < kact[..., None, None]) & (i[None, :] < lt[..., None, None]))
    R = torch.matmul(X.transpose(-1, -2).contiguous(), X)                     # [rx, d, L, L]
    Rs = torch.zeros_like(R)
    for s in range(-T, T + 1):
        lo, hi = max(0, -s), n_d - max(0, s)
        if hi > lo:
            Rs[:, lo:hi] += R[:, lo + s:hi + s]
    inb = i < lt[..., None]                                       # [rx, d, L]
    A = Rs * (inb[..., :, None] & inb[..., None, :])
    tr = torch.diagonal(A, dim1=-2, dim2=-1).sum(-1)
    delta = loading_factor * tr / lt
    diag = torch.where(inb, delta[..., None], torch.ones_like(delta)[..., None])
    A = A + torch.diag_embed(diag)
    rhs = inb.to(A.dtype)[..., None]
    w = torch.linalg.solve_ex(A, rhs)[0][..., 0]      # singular (all-zero) samples end as NaN -> 0
    wsum = w.sum(-1)
    xs = X.sum(2)
    y = (xs * w).sum(-1) / (wsum * kact.to(A.dtype) ** scaling_power)
    y = torch.where(torch.isnan(y) | (y.abs() > 1e7), torch.zeros_like(y), y) * n_ch
    return y.transpose(0, 1).contiguous()
