"""DAS on the GPU against the host oracle (bit-exact in FP32), through the
C ABI (supra_das_run / supra_das_run_frames via the RxBeamformer mirror)."""
import numpy as np
import pytest

from oracle import oracle as O
from supra_amd import _native as N
from supra_amd import workloads as W
from supra_amd.beamformer import BeamformingNode, RxBeamformer, USRawData
from supra_amd.geometry import BeamformerConfig, RxGeometry

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def gpu_das(geom, rf, fs, out_dtype=None, sample_beamformer=N.BF_DELAY_AND_SUM, **kw):
    raw = USRawData(torch.from_numpy(np.ascontiguousarray(rf)).cuda(), fs, geom)
    bf = RxBeamformer(geom)
    img = bf.perform_rx_beamforming(sample_beamformer, raw, kw["f_number"], kw["speed_of_sound_mm_per_s"],
                                    kw["window_type"], kw.get("window_parameter", 0.0),
                                    kw.get("interpolate_transmits", False), kw.get("additional_offset", 0),
                                    out_dtype=out_dtype)
    torch.cuda.synchronize()
    out = img.data.cpu().numpy()
    bf.close()
    return out


def oracle_das(geom, rf, fs, precision=32, **kw):
    return O.das(geom, rf, fs, precision=precision, **kw)


def assert_bit_exact(gpu, ref, label):
    same = (gpu == ref) | (np.isnan(gpu) & np.isnan(ref))
    if not same.all():
        bad = np.argwhere(~same)
        i = tuple(bad[0])
        pytest.fail(f"{label}: {len(bad)} of {gpu.size} pixels differ; first {i}: gpu {gpu[i]!r} oracle {ref[i]!r}")


def rel_err(a, ref):
    """max |a - ref| / max |ref| over pixels that saturate in neither (the
    single-precision 0/0 of an element-aligned depth-0 pixel clamps to
    -FLT_MAX; double precision does not see that element at depth 0)."""
    m = np.isfinite(ref) & (np.abs(ref) < 1e30) & (np.abs(a) < 1e30)
    den = np.abs(ref[m]).max()
    return float(np.abs(a[m].astype(np.float64) - ref[m]).max() / den)


def test_linear_2d_int16_bit_exact(device):
    w = W.linear_2d()
    g = w.geometry()
    rf = w.rf_host()
    kw = w.das_kwargs()
    gpu = gpu_das(g, rf, w.fs_hz, **kw)
    ref = oracle_das(g, rf, w.fs_hz, **kw)
    assert_bit_exact(gpu, ref, "linear_2d")
    ref64 = oracle_das(g, rf, w.fs_hz, precision=64, **kw)
    print(f"linear_2d: bit-exact vs FP32 oracle; normwise rel err vs FP64 oracle {rel_err(gpu, ref64):.3e}")


@pytest.mark.parametrize("window,param", [(N.WINDOW_RECTANGULAR, 0.0), (N.WINDOW_HAMMING, 0.1),
                                          (N.WINDOW_GAUSS, 0.4)])
def test_linear_2d_windows_float_rf(device, window, param):
    w = W.linear_2d(num_scatterers=40, f_number=1.7, window=window, window_parameter=param)
    g = w.geometry()
    rf = w.rf_host(dtype=np.float32)
    kw = w.das_kwargs()
    assert_bit_exact(gpu_das(g, rf, w.fs_hz, **kw), oracle_das(g, rf, w.fs_hz, **kw), f"window {window}")


def test_linear_2d_int16_output_and_offset(device):
    w = W.linear_2d(num_scatterers=30, additional_offset=-23)
    g = w.geometry()
    rf = w.rf_host()
    kw = w.das_kwargs()
    gpu = gpu_das(g, rf, w.fs_hz, out_dtype=torch.int16, **kw)
    ref = O.das(g, rf, w.fs_hz, out_int16=True, **kw).astype(np.int16)
    assert gpu.dtype == np.int16
    assert_bit_exact(gpu, ref, "int16 output, offset -23")
    kw["additional_offset"] = 41
    assert_bit_exact(gpu_das(g, rf, w.fs_hz, **kw), O.das(g, rf, w.fs_hz, **kw), "offset +41")


def test_interpolated_transmits_and_steering(device):
    cfg = BeamformerConfig(num_scanlines=(48, 1), element_layout=(96, 1), max_aperture=(48, 1),
                           tx_max_aperture=(48, 1), rx_scanline_subdivision=(3, 1), num_depths=700,
                           depth=30.0, steering_angle_deg=(8.0, 0.0), sector_angle_deg=(10.0, 0.0))
    w = W.linear_2d(bf=cfg, num_channels=96, num_samples=1600, num_scatterers=50)
    g = w.geometry()
    rf = w.rf_host()
    for interp in (False, True):
        kw = dict(w.das_kwargs(), interpolate_transmits=interp)
        assert_bit_exact(gpu_das(g, rf, w.fs_hz, **kw), O.das(g, rf, w.fs_hz, **kw), f"interp={interp}")


def test_sparse_element_map(device):
    w = W.linear_2d(num_scatterers=30)
    g = w.geometry()
    rng = np.random.default_rng(7)
    emap = (np.arange(128) % 64).astype(np.int32)
    emap[rng.choice(128, 20, replace=False)] = -1
    g.element_to_channel = emap
    rf = w.rf_host()
    kw = w.das_kwargs()
    assert_bit_exact(gpu_das(g, rf, w.fs_hz, **kw), O.das(g, rf, w.fs_hz, **kw), "sparse map")


def test_test_signal_known_answer(device):
    w = W.linear_2d()
    g = w.geometry()
    rf = np.zeros(w.rf_shape, dtype=np.int16)
    kw = w.das_kwargs()
    gpu = gpu_das(g, rf, w.fs_hz, sample_beamformer=N.BF_TEST_SIGNAL, **kw)
    ref = O.das(g, rf, w.fs_hz, test_signal=True, **kw)
    assert_bit_exact(gpu, ref, "test signal")
    assert set(np.unique(gpu)) == {0.0, 1000.0}


def test_matrix_3d_small_bit_exact(device):
    w = W.matrix_3d_small(n_el=12, n_lines=10, n_depths=300)
    g = w.geometry()
    rf = w.rf_host()
    kw = w.das_kwargs()
    gpu = gpu_das(g, rf, w.fs_hz, **kw)
    assert_bit_exact(gpu, O.das(g, rf, w.fs_hz, **kw), "matrix_3d small")
    ref64 = O.das(g, rf, w.fs_hz, precision=64, **kw)
    print(f"matrix_3d small: normwise rel err vs FP64 {rel_err(gpu, ref64):.3e}")
    kw["additional_offset"] = -7
    assert_bit_exact(gpu_das(g, rf, w.fs_hz, **kw), O.das(g, rf, w.fs_hz, **kw), "matrix_3d offset -7")
    gts = gpu_das(g, rf, w.fs_hz, sample_beamformer=N.BF_TEST_SIGNAL, **kw)
    assert_bit_exact(gts, O.das(g, rf, w.fs_hz, test_signal=True, **kw), "matrix_3d test signal")


def test_frames_batch_matches_single(device):
    w = W.linear_2d(num_scatterers=20)
    g = w.geometry()
    rfs = np.stack([w.rf_host(), W.linear_2d(num_scatterers=25, scatter_seed=9, noise_seed=10).rf_host()])
    kw = w.das_kwargs()
    batch = gpu_das(g, rfs, w.fs_hz, **kw)
    assert batch.shape == (2, 2080, 128)
    for f in range(2):
        assert_bit_exact(batch[f], O.das(g, rfs[f], w.fs_hz, **kw), f"frame {f}")


@pytest.mark.parametrize("n_frames,rf_dt,out_dt,sparse", [(3, "i16", "f32", False), (8, "i16", "i16", False),
                                                        (11, "f32", "f32", False), (9, "i16", "f32", True)])
def test_frames_per_wave_bit_exact(device, n_frames, rf_dt, out_dt, sparse):
    """Frames that share a geometry run several per wave (one delay
    evaluation per element and pixel for all of them): full and partial frame
    groups, int16 and float RF and images, and a sparse element map are bit
    for bit the oracle's, frame by frame."""
    w = W.linear_2d(num_scatterers=15)
    g = w.geometry()
    if sparse:
        emap = (np.arange(128) % 64).astype(np.int32)
        emap[np.random.default_rng(3).choice(128, 17, replace=False)] = -1
        g.element_to_channel = emap
    dt = np.int16 if rf_dt == "i16" else np.float32
    rfs = np.stack([W.linear_2d(num_scatterers=15, scatter_seed=40 + f, noise_seed=60 + f).rf_host(dtype=dt)
                    for f in range(n_frames)])
    kw = w.das_kwargs()
    odt = torch.int16 if out_dt == "i16" else None
    batch = gpu_das(g, rfs, w.fs_hz, out_dtype=odt, **kw)
    assert batch.shape == (n_frames, 2080, 128)
    for f in range(n_frames):
        ref = O.das(g, rfs[f], w.fs_hz, out_int16=(out_dt == "i16"), **kw)
        if out_dt == "i16":
            ref = ref.astype(np.int16)
        assert_bit_exact(batch[f], ref, f"frame {f} of {n_frames}")


def test_node_dispatch_and_xml(device):
    xml = """<node type="BeamformingNode" id="BEAM">
      <param name="fNumber" type="double">1</param>
      <param name="windowType" type="string">Hann</param>
      <param name="speedOfSound" type="double">1540</param>
      <param name="beamformerType" type="string">DelayAndSum</param>
      <param name="outputType" type="string">float</param>
    </node>"""
    node = BeamformingNode.from_xml(xml)
    w = W.linear_2d(num_scatterers=20)
    g = w.geometry()
    rf = w.rf_host()
    img = node.process(USRawData(torch.from_numpy(rf).cuda(), w.fs_hz, g))
    torch.cuda.synchronize()
    assert_bit_exact(img.data.cpu().numpy(), O.das(g, rf, w.fs_hz, **w.das_kwargs()), "node")
    assert img.properties["Beamformer.WindowType"] == "Hann"


def test_rejects_out_of_range_transmit(device):
    w = W.linear_2d()
    g = w.geometry()
    rf = torch.zeros((64, 64, 2080), dtype=torch.int16, device="cuda")   # fewer transmits than referenced
    bf = RxBeamformer(g)
    with pytest.raises(N.SupraError):
        bf.perform_rx_beamforming(0, USRawData(rf, w.fs_hz, g), 1.0, 1.54e6, 1, 0.0, False, 0)
    bf.close()


@pytest.mark.parametrize("num_rx,num_depths,aperture", [(8, 2, 8), (16, 63, 16), (16, 65, 8), (24, 257, 16),
                                                        (64, 1000, 32)])
def test_ragged_sizes_bit_exact(device, num_rx, num_depths, aperture):
    """Depth counts that are not multiples of the 64-depth wave unit (down to
    the geometry's minimum of two), small arrays and apertures: every pixel is
    written and matches the oracle bit for bit."""
    depth_mm = num_depths / 40e6 * 1540e3 / 2.0
    w = W.linear_acquisition(num_rx, num_depths, aperture, 40e6, 20, scatter_seed=7, noise_seed=8,
                             box_depth=(0.0, depth_mm))
    g = w.geometry()
    rf = w.rf_host()
    kw = w.das_kwargs()
    gpu = gpu_das(g, rf, w.fs_hz, **kw)
    assert gpu.shape == (num_depths, num_rx)
    assert_bit_exact(gpu, oracle_das(g, rf, w.fs_hz, **kw), f"{num_rx}x{num_depths} aperture {aperture}")


@pytest.mark.parametrize("num_rx,num_depths,n_frames", [(40, 130, 3), (7, 65, 2), (33, 257, 9), (64, 2, 2)])
def test_frames_ragged_sizes_bit_exact(device, num_rx, num_depths, n_frames):
    """Several frames per launch (the multi-frame kernel, tiled by 32
    scanlines x 2 depth chunks) at sizes that leave partial tiles: scanline
    counts off the 32 grid, odd and even chunk counts, a partial frame group.
    Every frame matches the single-frame oracle bit for bit."""
    depth_mm = num_depths / 40e6 * 1540e3 / 2.0
    w = W.linear_acquisition(num_rx, num_depths, 16, 40e6, 20, scatter_seed=9, noise_seed=10,
                             box_depth=(0.0, depth_mm))
    g = w.geometry()
    kw = w.das_kwargs()
    frames = [W.linear_acquisition(num_rx, num_depths, 16, 40e6, 20, scatter_seed=9 + f, noise_seed=10 + f,
                                   box_depth=(0.0, depth_mm)).rf_host() for f in range(n_frames)]
    batch = gpu_das(g, np.stack(frames), w.fs_hz, **kw)
    assert batch.shape == (n_frames, num_depths, num_rx)
    for f in range(n_frames):
        assert_bit_exact(batch[f], oracle_das(g, frames[f], w.fs_hz, **kw),
                         f"frame {f} of {n_frames}, {num_rx}x{num_depths}")


@pytest.mark.parametrize("n_depths", [2, 65, 130])
def test_matrix_3d_ragged_depths_bit_exact(device, n_depths):
    w = W.matrix_3d_small(n_el=4, n_lines=4, n_depths=n_depths, n_samples=600)
    g = w.geometry()
    rf = w.rf_host()
    kw = w.das_kwargs()
    assert_bit_exact(gpu_das(g, rf, w.fs_hz, **kw), oracle_das(g, rf, w.fs_hz, **kw), f"3D {n_depths} depths")
