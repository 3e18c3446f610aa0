"""Differentiable 2D delay-and-sum written with plain PyTorch operations
(gather, elementwise arithmetic, a sum over the aperture) and torch autograd,
one frame at a time: the plain per-frame GPU form of the operation that the
hand-written forward/adjoint kernels are compared with, for accuracy in the
GPU tests and for speed in bench.py.

TEST AND BENCHMARK INFRASTRUCTURE: never imported by the supra_amd package.

Everything that does not depend on the speed of sound is tabulated once from
the millimetre geometry in float64: aperture membership |x_e - x_s| <= d/(2F)
(RxSampleBeamformerDelayAndSum.h:167), the window weight
(WindowFunction.h:58-64), n/W, the receive path length in millimetres.  Per
frame, tau = t0 + path * fs / c + offset is formed in float32 with c a
tensor, so autograd differentiates the linear interpolation with respect to
the RF (through the gather) and to c.  The highest-weight transmit only (no
transmit interpolation); the i == N_s - 1, f == 0 edge sample is dropped.
"""
from __future__ import annotations

import numpy as np
import torch

from supra_amd.beamformer import window_lut


class PlainTorchDAS:
    def __init__(self, geometry, fs_hz, f_number, window_type, window_parameter, num_channels, num_samples,
                 additional_offset=0, device="cuda"):
        g = geometry
        self.n_rx, self.n_d, self.n_s = g.num_rx, g.num_depths, num_samples
        self.fs = float(fs_hz)
        self.offset = float(additional_offset)
        lut = window_lut(window_type, window_parameter).astype(np.float64)
        sc = g.scanlines
        sel = np.argmax(sc["tx_weights"], axis=1)
        rows = np.arange(self.n_rx)
        tx = sc["tx"][rows, sel]
        first = tx["first"][:, 0].astype(np.int64)
        last = tx["last"][:, 0].astype(np.int64)
        A = int((last - first).max())
        e = first[:, None] + np.arange(A)[None, :]                     # [rx, A]
        inap = e < last[:, None]
        e = np.minimum(e, g.num_elements - 1)
        xe = g.element_x.astype(np.float64)[e]
        xs = sc["position"][:, 0]
        dirx, diry = sc["direction"][:, 0], sc["direction"][:, 1]
        med = sc["max_element_distance"][:, 0]
        d = g.depths.astype(np.float64)                               # [D]
        dx = xe - xs[:, None]                                         # [rx, A]
        half = d / (2 * f_number)                                     # [D]
        member = (np.abs(dx)[:, None, :] <= half[None, :, None]) & inap[:, None, :]
        with np.errstate(divide="ignore", invalid="ignore"):
            rel = dx[:, None, :] / np.minimum(half[None, :, None], med[:, None, None])
        rel = np.where(np.isnan(rel), -1.0, np.clip(rel, -1.0, 1.0))
        w = np.where(member, lut[np.floor(31.5 * (rel + 1.0) + 0.5).astype(np.int64)], 0.0)
        n = member.sum(-1)
        W = w.sum(-1)
        with np.errstate(divide="ignore", invalid="ignore"):
            norm = np.where((n > 0) & (W > 0), n / W, 0.0)
        path = np.sqrt((xe[:, None, :] - (xs[:, None, None] + dirx[:, None, None] * d[None, :, None])) ** 2 +
                       (diry[:, None, None] * d[None, :, None]) ** 2) + d[None, :, None]
        ch = e % num_channels
        line = (tx["tx_idx"].astype(np.int64)[:, None] * num_channels + ch) * num_samples   # [rx, A]
        dev = torch.device(device)
        self.wn = torch.tensor(w * norm[..., None], dtype=torch.float32, device=dev)
        self.path = torch.tensor(path, dtype=torch.float32, device=dev)
        self.t0 = torch.tensor(tx["initial_delay"] * fs_hz, dtype=torch.float32, device=dev)
        self.line = torch.tensor(line, dtype=torch.int64, device=dev)

    def __call__(self, rf_frame, c_m_per_s):
        """rf_frame [tx][ch][s] float32, c a 0-d tensor in m/s -> image [depth][scanline]."""
        scale = self.fs / (c_m_per_s.to(torch.float32) * 1000.0)
        tau = self.t0[:, None, None] + self.path * scale + self.offset
        i = torch.floor(tau).detach()
        f = tau - i
        ii = i.to(torch.int64)
        valid = (ii >= 0) & (ii < self.n_s - 1)
        idx = self.line[:, None, :] + ii.clamp(0, self.n_s - 2)
        flat = rf_frame.reshape(-1)
        v0 = flat[idx]
        v1 = flat[idx + 1]
        val = ((1.0 - f) * v0 + f * v1) * valid
        return (self.wn * val).sum(-1).transpose(0, 1)

    def delayed(self, rf_frame, c_m_per_s, num_channels):
        """The pre-delay of RawDelay (RawDelay.cu:221-312) in the same plain
        form: per aperture element j = e - first, w_e (n/W) interp(tau)
        -> [scanline][num_channels][depth] (channels beyond the aperture 0)."""
        scale = self.fs / (c_m_per_s.to(torch.float32) * 1000.0)
        tau = self.t0[:, None, None] + self.path * scale
        i = torch.floor(tau)
        f = tau - i
        ii = i.to(torch.int64)
        valid = (ii >= 0) & (ii < self.n_s - 1)
        idx = self.line[:, None, :] + ii.clamp(0, self.n_s - 2)
        flat = rf_frame.reshape(-1)
        val = ((1.0 - f) * flat[idx] + f * flat[idx + 1]) * valid
        d = (self.wn * val).permute(0, 2, 1)                           # [rx][A][D]
        out = torch.zeros((self.n_rx, num_channels, self.n_d), dtype=torch.float32, device=d.device)
        a = min(d.shape[1], num_channels)
        out[:, :a] = d[:, :a]
        return out
