"""Pre-delay (bit-exact vs the FP32 oracle) and minimum variance (vs the FP64
oracle, per-sample relative error) on the GPU through the C ABI."""
import numpy as np
import pytest

from oracle import oracle as O
from supra_amd import _native as N
from supra_amd import workloads as W
from supra_amd.beamformer import (BeamformingMVNode, RawDelay, RawDelayNode, USRawData,
                                  mv_perform_rx_beamforming)

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def gpu_mv(x, **kw):
    raw = USRawData(torch.from_numpy(np.ascontiguousarray(x)).cuda(), 1.0, None, 1, (1, 1))
    img = mv_perform_rx_beamforming(raw, **kw)
    torch.cuda.synchronize()
    return img.data.cpu().numpy()


def rel_errors(gpu, ref, floor_frac=1e-3):
    den = np.maximum(np.abs(ref), floor_frac * np.abs(ref).max())
    return np.abs(gpu.astype(np.float64) - ref) / den


def test_rawdelay_bit_exact(device):
    w = W.linear_2d()
    g = w.geometry()
    rf = w.rf_host()
    rd = RawDelay(g)
    for win, odt in [(N.WINDOW_HANN, torch.float32), (N.WINDOW_HAMMING, torch.int16)]:
        res = rd.perform_delay(USRawData(torch.from_numpy(rf).cuda(), w.fs_hz, g), 1.0, 1.54e6, win, 0.0,
                               out_dtype=odt)
        torch.cuda.synchronize()
        gpu = res.data.cpu().numpy()
        ref = O.rawdelay(g, rf, w.fs_hz, 1.0, 1.54e6, win, out_int16=(odt == torch.int16))
        if odt == torch.int16:
            ref = ref.astype(np.int16)
        same = (gpu == ref) | (np.isnan(gpu) & np.isnan(ref))
        assert same.all(), f"{(~same).sum()} pre-delayed samples differ (window {win})"
        assert abs(res.sampling_frequency - 1 / (40.0 / 1.54e6 / 2080)) < 1e-6
    rd.close()


@pytest.mark.parametrize("n_el,odt", [(6, torch.float32), (8, torch.int16)])
def test_rawdelay3d_bit_exact(device, n_el, odt):
    """3D pre-delay of a matrix probe (RawDelay.cu:126-220, 315-402): every
    slot of every pixel bit for bit the FP32 oracle's, then minimum variance
    on the pre-delayed volume against the FP64 oracle."""
    w = W.matrix_3d_small(n_el=n_el, n_lines=5, n_depths=150, n_samples=700)
    g = w.geometry()
    rf = w.rf_host()
    rd = RawDelay(g)
    res = rd.perform_delay(USRawData(torch.from_numpy(rf).cuda(), w.fs_hz, g), 1.0, 1.54e6, N.WINDOW_HANN, 0.0,
                           out_dtype=odt)
    torch.cuda.synchronize()
    gpu = res.data.cpu().numpy()
    ref = O.rawdelay(g, rf, w.fs_hz, 1.0, 1.54e6, N.WINDOW_HANN, out_int16=(odt == torch.int16))
    if odt == torch.int16:
        ref = ref.astype(np.int16)
    assert gpu.shape == (g.num_rx, n_el * n_el, g.num_depths)
    same = (gpu == ref) | (np.isnan(gpu) & np.isnan(ref))
    assert same.all(), f"{(~same).sum()} pre-delayed samples differ"
    if odt == torch.float32:
        L = n_el * n_el // 2
        img = mv_perform_rx_beamforming(res, L, 2, 1.5, False, 0.05)
        torch.cuda.synchronize()
        out = img.data.cpu().numpy()
        mref = O.mv(gpu, L, 2, 1.5, False, 0.05)
        huge = (np.abs(gpu) > 1e19).any(axis=1)
        keep = ~np.array([np.convolve(h, np.ones(5), "same") > 0 for h in huge]).T
        err = rel_errors(out[keep], mref[keep])
        print(f"MV on the 3D pre-delayed volume: max {err.max():.3e} p99 {np.quantile(err, 0.99):.3e}")
        assert np.quantile(err, 0.99) < 1e-4 and err.max() < 1e-2
    rd.close()


def test_rawdelay_node(device):
    w = W.linear_2d(num_scatterers=30)
    g = w.geometry()
    rf = w.rf_host()
    node = RawDelayNode(fNumber=1.0, windowType="Hann")
    res = node.process(USRawData(torch.from_numpy(rf).cuda(), w.fs_hz, g))
    torch.cuda.synchronize()
    ref = O.rawdelay(g, rf, w.fs_hz, 1.0, 1.54e6, N.WINDOW_HANN)
    assert np.array_equal(res.data.cpu().numpy(), ref)


@pytest.mark.parametrize("alpha,T,p", [(0.01, 5, 1.5), (1.0, 0, 1.0)])
def test_mv_plane_wave_closed_form(device, alpha, T, p):
    rng = np.random.default_rng(4)
    n_rx, n_ch, n_d = 16, 64, 300
    s = (rng.uniform(1.0, 3.0, (n_rx, 1, n_d)) * rng.choice([-1.0, 1.0], (n_rx, 1, n_d))).astype(np.float32)
    x = np.ascontiguousarray(np.broadcast_to(s, (n_rx, n_ch, n_d)))
    out = gpu_mv(x, sub_array_size=32, temporal_smoothing=T, sub_array_scaling_power=p, loading_factor=alpha)
    expected = (n_ch * s[:, 0, :].astype(np.float64) * 33 ** (1 - p)).T
    assert np.abs(out / expected - 1).max() < 1e-4


@pytest.mark.parametrize("alpha,T,p", [(0.01, 5, 1.5), (1.0, 2, 1.0)])
def test_mv_means_channel_ramp_closed_form(device, alpha, T, p):
    """Mean removal on x[c, t] = a(t) + b(t) c (exact in float32): R~ is
    rank one along 1 and y = K^(1-p) (a + b ((K-1)/2 + (L-1)/2)), times N_ch,
    independent of alpha and T."""
    rng = np.random.default_rng(13)
    n_rx, n_ch, n_d, L = 8, 64, 200, 32
    K = n_ch - L + 1
    a = rng.integers(200, 300, (n_rx, n_d)).astype(np.float64)
    b = rng.integers(4, 9, (n_rx, n_d)) / 4.0 * rng.choice([-1.0, 1.0], (n_rx, n_d))
    x = (a[:, None, :] + b[:, None, :] * np.arange(n_ch)[None, :, None]).astype(np.float32)
    out = gpu_mv(x, sub_array_size=L, temporal_smoothing=T, sub_array_scaling_power=p, compute_means=True,
                 loading_factor=alpha)
    expected = (n_ch * K ** (1 - p) * (a + b * ((K - 1) / 2 + (L - 1) / 2))).T
    err = np.abs(out / expected - 1).max()
    print(f"MV means ramp closed form alpha={alpha} T={T}: max rel err {err:.2e}")
    assert err < 1e-4


def test_mv_predelayed_frame_vs_fp64_oracle(device):
    """The minimum-variance workload: pre-delayed linear_2d frame, L = 32, T = 5,
    loading trace/(100 L)."""
    w = W.linear_2d()
    g = w.geometry()
    rf = w.rf_host()
    rd = RawDelay(g)
    x = rd.perform_delay(USRawData(torch.from_numpy(rf).cuda(), w.fs_hz, g), 1.0, 1.54e6, N.WINDOW_HANN, 0.0)
    img = mv_perform_rx_beamforming(x, 32, 5, 1.5, False, 0.01)
    torch.cuda.synchronize()
    gpu = img.data.cpu().numpy()
    xh = x.data.cpu().numpy()
    ref = O.mv(xh, 32, 5, 1.5, False, 0.01)
    # Pre-delayed samples at -FLT_MAX (the reference's clampCast of the 0/0 of
    # a Hann window at depth 0) overflow FP32 covariances; those samples end
    # as NaN -> 0 in single precision (as in the reference), not in FP64.
    huge = (np.abs(xh) > 1e19).any(axis=1)                    # [scanline, sample]
    T = 5
    tainted = np.zeros_like(huge)
    for d in range(-T, T + 1):
        tainted |= np.roll(huge, d, axis=1) if d else huge
    tainted = tainted.T
    assert np.all(gpu[tainted] == 0)
    err = rel_errors(gpu[~tainted], ref[~tainted])
    print(f"MV per-sample relative error vs FP64 oracle ({(~tainted).sum()} samples): max {err.max():.3e} "
          f"p99 {np.quantile(err, 0.99):.3e} median {np.median(err):.3e}")
    assert np.quantile(err, 0.99) < 1e-4
    assert err.max() < 1e-2
    rd.close()


@pytest.mark.parametrize("L,T,means,n_ch", [(16, 2, False, 48), (32, 3, True, 64), (64, 5, False, 128),
                                            (8, 0, False, 32)])
def test_mv_random_apertures(device, L, T, means, n_ch):
    rng = np.random.default_rng(L + T)
    n_rx, n_d = 6, 257
    x = rng.normal(size=(n_rx, n_ch, n_d)).astype(np.float32) * 100
    f = rng.integers(0, n_ch // 3, (n_rx, n_d))
    l = np.minimum(n_ch - 1, f + rng.integers(3, n_ch, (n_rx, n_d)))
    ch = np.arange(n_ch)[None, :, None]
    x[(ch < f[:, None, :]) | (ch > l[:, None, :])] = 0.0
    x[:, :, 7] = 0.0                      # an all-zero sample
    gpu = gpu_mv(x, sub_array_size=L, temporal_smoothing=T, compute_means=means, loading_factor=0.05)
    ref = O.mv(x, L, T, 1.5, means, 0.05)
    err = rel_errors(gpu, ref)
    print(f"L={L} T={T} means={means}: max {err.max():.3e} p99 {np.quantile(err, 0.99):.3e}")
    assert np.quantile(err, 0.99) < 1e-4 and err.max() < 1e-2
    assert gpu[7].tolist() == ref[7].tolist()


def test_mv_int16_io_and_node(device):
    rng = np.random.default_rng(9)
    x = (rng.normal(size=(4, 64, 200)) * 300).astype(np.int16)
    node = BeamformingMVNode(subArraySize=32, temporalSmoothing=3, outputType="int16", loadingFactor=0.1)
    img = node.process(USRawData(torch.from_numpy(x).cuda(), 1.0, None, 1, (1, 1)))
    torch.cuda.synchronize()
    gpu = img.data.cpu().numpy()
    ref = O.mv(x, 32, 3, 1.5, False, 0.1, out_int16=True)
    assert gpu.dtype == np.int16
    assert np.abs(gpu.astype(np.int64) - ref.astype(np.int64)).max() <= 1


def test_mv_rejects_large_subarray(device):
    x = torch.zeros((2, 160, 50), device="cuda")
    with pytest.raises(N.SupraError):
        mv_perform_rx_beamforming(USRawData(x, 1.0, None, 1, (1, 1)), sub_array_size=80)


def test_mv_vs_plain_torch(device):
    """The fused kernel against the same algorithm written in plain PyTorch
    (float32, torch.linalg.solve) on a pre-delayed frame."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).parent))
    from torch_mv_reference import plain_mv

    w = W.linear_2d(num_scatterers=150)
    g = w.geometry()
    rf = w.rf_host()
    rd = RawDelay(g)
    x = rd.perform_delay(USRawData(torch.from_numpy(rf).cuda(), w.fs_hz, g), 1.0, 1.54e6, N.WINDOW_HANN, 0.0)
    img = mv_perform_rx_beamforming(x, 32, 5, 1.5, False, 0.01)
    ref = plain_mv(x.data, 32, 5, 1.5, 0.01)
    torch.cuda.synchronize()
    gpu, ref = img.data.cpu().numpy(), ref.cpu().numpy().astype(np.float64)
    huge = (np.abs(x.data.cpu().numpy()) > 1e19).any(axis=1)
    tainted = np.zeros_like(huge)
    for d in range(-5, 6):
        tainted |= np.roll(huge, d, axis=1)
    keep = ~tainted.T
    err = rel_errors(gpu[keep], ref[keep])
    print(f"MV kernel vs plain PyTorch float32: max {err.max():.3e} p99 {np.quantile(err, 0.99):.3e}")
    assert np.quantile(err, 0.99) < 1e-3
    rd.close()


@pytest.mark.parametrize("win,L,T,means,odt", [(N.WINDOW_HANN, 32, 5, False, "f32"),
                                               (N.WINDOW_HAMMING, 16, 3, True, "i16")])
def test_fused_predelay_mv_matches_two_step(device, win, L, T, means, odt):
    """supra_mv_das_run equals supra_rawdelay_run (float) + supra_mv_run bit for bit."""
    from supra_amd.beamformer import RxBeamformer, mv_das_perform_rx_beamforming
    w = W.linear_2d(num_scatterers=120)
    g = w.geometry()
    rf = torch.from_numpy(w.rf_host()).cuda()
    out_dtype = torch.float32 if odt == "f32" else torch.int16
    rd = RawDelay(g)
    x = rd.perform_delay(USRawData(rf, w.fs_hz, g), 1.0, 1.54e6, win, 0.0)
    two = mv_perform_rx_beamforming(x, L, T, 1.5, means, 0.01, out_dtype=out_dtype)
    eng = RxBeamformer(g)
    one = mv_das_perform_rx_beamforming(eng, USRawData(rf, w.fs_hz, g), 1.0, 1.54e6, win, 0.0, L, T, 1.5, means,
                                        0.01, out_dtype=out_dtype)
    torch.cuda.synchronize()
    a, b = one.data.cpu().numpy(), two.data.cpu().numpy()
    assert np.array_equal(a, b), f"{(a != b).sum()} samples differ"
    rd.close()
    eng.close()


def test_plain_torch_rawdelay_baseline(device):
    """The plain PyTorch pre-delay that bench.py times as the per-frame GPU
    form does the same work as the kernel: same layout and values up to its
    float64 geometry tables (aperture-edge ties) and the dropped last-sample
    edge case."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parent))
    from torch_das_reference import PlainTorchDAS

    w = W.linear_2d(num_scatterers=50)
    g = w.geometry()
    rf = torch.from_numpy(w.rf_host()).cuda()
    rd = RawDelay(g)
    ours = rd.perform_delay(USRawData(rf, w.fs_hz, g), 1.0, 1.54e6, N.WINDOW_HANN, 0.0).data
    plain = PlainTorchDAS(g, w.fs_hz, 1.0, N.WINDOW_HANN, 0.0, w.num_channels, w.rf_shape[2])
    theirs = plain.delayed(rf.to(torch.float32), torch.tensor(1540.0, device="cuda"), w.num_channels)
    torch.cuda.synchronize()
    a, b = ours.cpu().numpy().astype(np.float64), theirs.cpu().numpy().astype(np.float64)
    a = np.where(np.abs(a) > 1e30, 0.0, a)          # clampCast of 0/0 at depth 0 (-FLT_MAX)
    assert a.shape == b.shape
    assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-2
    rd.close()


@pytest.mark.parametrize("n_rx,n_d,L,T", [(1, 1, 8, 5), (3, 2, 16, 5), (2, 11, 32, 5), (5, 65, 64, 3)])
def test_mv_ragged_sizes(device, n_rx, n_d, L, T):
    """Depth counts below, at and just past the temporal window and the depth
    tile, one scanline, and the largest subarray: every sample matches the
    FP64 oracle."""
    rng = np.random.default_rng(n_d + L)
    n_ch = max(L, 16) + 8
    x = rng.normal(size=(n_rx, n_ch, n_d)).astype(np.float32) * 50
    gpu = gpu_mv(x, sub_array_size=L, temporal_smoothing=T, loading_factor=0.05)
    ref = O.mv(x, L, T, 1.5, False, 0.05)
    assert gpu.shape == ref.shape == (n_d, n_rx)
    err = rel_errors(gpu, ref)
    assert np.quantile(err, 0.99) < 1e-4 and err.max() < 1e-2
