"""supra_das_forward_c_noalloc through the C ABI: the same images and
normalisation as supra_das_forward_c (bit for bit) on two calls in a row, the
second with a smaller input, and a clear error for tables beyond the
file-scope scratch."""
import ctypes as C

import numpy as np
import pytest

from supra_amd import _native as N
from supra_amd import workloads as W
from supra_amd.beamformer import RxBeamformer

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def vp(t):
    return C.c_void_p(t.data_ptr())


def forward(entry, eng, x, c_mm, stream=0):
    lib = N.lib()
    frames, n_tx, n_ch, n_s = (int(v) for v in x.shape)
    g = eng.geometry
    out = torch.full((frames, g.num_depths, g.num_rx), 7.0, dtype=torch.float32, device="cuda")
    norm = torch.full((frames, 1, g.num_rx, g.num_depths), 7.0, dtype=torch.float32, device="cuda")
    code = getattr(lib, entry)(eng._ctx, frames, vp(x), n_tx, n_ch, n_s, 40e6, 1.0, vp(c_mm), N.WINDOW_HANN, 0.0,
                               0, 0, vp(out), vp(norm), C.c_void_p(stream))
    return code, out, norm


def test_noalloc_forward_matches_allocating_entry_point(device):
    cases = [(W.linear_2d(num_scatterers=40), 6), (W.linear_acquisition(64, 900, 32, 40e6, 30, 3, 4), 2)]
    for step, (w, frames) in enumerate(cases):          # second call: fewer frames, a smaller image
        g = w.geometry()
        eng = RxBeamformer(g)
        rf = np.stack([w.rf_host(dtype=np.float32) * (1.0 + 0.1 * f) for f in range(frames)]).astype(np.float32)
        x = torch.from_numpy(rf).cuda()
        c_mm = torch.tensor(np.linspace(1460e3, 1590e3, frames), dtype=torch.float64, device="cuda")
        code_a, out_a, norm_a = forward("supra_das_forward_c", eng, x, c_mm)
        code_b, out_b, norm_b = forward("supra_das_forward_c_noalloc", eng, x, c_mm)
        torch.cuda.synchronize()
        assert code_a == 0 and code_b == 0, N.lib().supra_last_error()
        a, b = out_a.cpu().numpy(), out_b.cpu().numpy()
        same = (a == b) | (np.isnan(a) & np.isnan(b))
        assert same.all(), f"call {step}: {(~same).sum()} pixels differ"
        assert np.array_equal(norm_a.cpu().numpy(), norm_b.cpu().numpy())
        eng.close()


def test_noalloc_forward_rejects_tables_beyond_scratch(device):
    w = W.linear_2d(num_scatterers=5)
    g = w.geometry()
    eng = RxBeamformer(g)
    per_frame = g.num_rx * 192 + g.num_depths * 4 + g.num_elements * 4 + 4
    frames = N.NOALLOC_SCRATCH_BYTES // per_frame + 2
    x = torch.zeros((frames, 1, 1, 1), dtype=torch.float32, device="cuda")   # never read: rejected first
    c_mm = torch.full((frames,), 1540e3, dtype=torch.float64, device="cuda")
    out = torch.zeros(1, dtype=torch.float32, device="cuda")
    code = N.lib().supra_das_forward_c_noalloc(eng._ctx, frames, vp(x), 128, 64, 2080, 40e6, 1.0, vp(c_mm),
                                               N.WINDOW_HANN, 0.0, 0, 0, vp(out), None, None)
    assert code == N.SUPRA_ERR_TOO_LARGE
    assert b"scratch" in N.lib().supra_last_error()
    eng.close()
