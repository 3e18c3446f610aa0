"""Envelope detection and log compression on the GPU through the C ABI:
the envelope bit for bit the FP32 oracle's (int16/float in and out, odd and
even filter lengths, images shorter than the filter), log compression within
the single-precision log10 of the two sides (uint8 within one grey level),
the fused launch bit for bit the two-step path, and the whole B-mode chain on
a beamformed linear_2d frame."""
import numpy as np
import pytest

from oracle import oracle as O
from supra_amd import _native as N
from supra_amd import workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def gpu_frame():
    from supra_amd.beamformer import RxBeamformer, USRawData
    w = W.linear_2d(num_scatterers=80)
    g = w.geometry()
    rf = torch.from_numpy(w.rf_host()).cuda()
    bf = RxBeamformer(g)
    kw = w.das_kwargs()
    img = bf.perform_rx_beamforming(N.BF_DELAY_AND_SUM, USRawData(rf, w.fs_hz, g), kw["f_number"],
                                    kw["speed_of_sound_mm_per_s"], kw["window_type"], 0.0, False, 0)
    torch.cuda.synchronize()
    bf.close()
    return img


@pytest.mark.parametrize("n_taps,in_dt,out_dt,shape", [(65, "f32", "f32", None), (65, "i16", "f32", (300, 70)),
                                                        (64, "f32", "i16", (257, 130)), (257, "f32", "f32", (90, 5)),
                                                        (31, "i16", "i16", (1, 3))])
def test_envelope_bit_exact(device, n_taps, in_dt, out_dt, shape):
    from supra_amd.postprocess import HilbertFirEnvelope
    if shape is None:
        img = gpu_frame().data
        img = torch.where(img.abs() < 1e30, img, torch.zeros_like(img))
    else:
        rng = np.random.default_rng(n_taps)
        x = rng.normal(size=shape) * 3000
        img = torch.from_numpy(x.astype(np.int16 if in_dt == "i16" else np.float32)).cuda()
    out_t = torch.int16 if out_dt == "i16" else torch.float32
    env = HilbertFirEnvelope(n_taps).demodulate(img, out_t)
    torch.cuda.synchronize()
    got = env.data.cpu().numpy()
    ref = O.clamp_cast(O.envelope(img.cpu().numpy(), n_taps), np.int16 if out_dt == "i16" else np.float32)
    assert got.dtype == ref.dtype and got.shape == ref.shape
    assert np.array_equal(got, ref), f"{(got != ref).sum()} of {got.size} envelope samples differ"
    assert env.properties["imageState"] == "EnvDetected"


@pytest.mark.parametrize("in_dt,out_dt,db,gain,in_max", [("f32", "f32", 80.0, 1.0, 1024.0),
                                                         ("i16", "u8", 60.0, 0.7, 3000.0),
                                                         ("u8", "f32", 40.0, 1.5, 200.0),
                                                         ("u8", "u8", 80.0, 1.0, 1024.0)])
def test_logcompress_vs_oracle(device, in_dt, out_dt, db, gain, in_max):
    from supra_amd.postprocess import LogCompressor
    rng = np.random.default_rng(11)
    dt = {"f32": np.float32, "i16": np.int16, "u8": np.uint8}[in_dt]
    if in_dt == "u8":
        x = rng.integers(0, 256, (123, 77)).astype(np.uint8)
    else:
        x = (rng.normal(size=(123, 77)) * 2000).astype(dt)
    out_t = torch.uint8 if out_dt == "u8" else torch.float32
    res = LogCompressor().compress(torch.from_numpy(x).cuda(), db, gain, in_max, out_t)
    torch.cuda.synchronize()
    got = res.data.cpu().numpy()
    ref32 = O.logcompress(x, db, gain, in_max)
    if out_dt == "u8":
        ref = O.clamp_cast(ref32, np.uint8)
        assert np.abs(got.astype(np.int32) - ref.astype(np.int32)).max() <= 1
        assert (got == ref).mean() > 0.99
    else:
        fin = np.isfinite(ref32)
        assert np.array_equal(np.isfinite(got), fin)
        assert np.abs(got[fin] - ref32[fin]).max() <= 4e-7 * max(1.0, np.abs(ref32[fin]).max())


@pytest.mark.parametrize("out_dt", ["u8", "f32"])
def test_bmode_chain_fused_equals_two_step(device, out_dt):
    """DAS -> envelope -> log compression: the fused launch is the two-step
    path bit for bit, and the chain agrees with the oracle (envelope exact,
    log compression within one grey level / 4e-7)."""
    from supra_amd.postprocess import (HilbertFirEnvelopeNode, LogCompressorNode, envelope_log_compress)
    img = gpu_frame()
    out_t = torch.uint8 if out_dt == "u8" else torch.float32
    env = HilbertFirEnvelopeNode(filterLength=65, outputType="float").process(img)
    two = LogCompressorNode(dynamicRange=60.0, gain=1.0, inMax=20000.0,
                            outputType="uint8" if out_dt == "u8" else "float").process(env)
    one = envelope_log_compress(img, 65, 60.0, 1.0, 20000.0, out_dtype=out_t)
    torch.cuda.synchronize()
    a, b = one.data.cpu().numpy(), two.data.cpu().numpy()
    assert np.array_equal(a, b), f"{(a != b).sum()} samples differ between fused and two-step"
    env_ref = O.envelope(img.data.cpu().numpy(), 65)
    assert np.array_equal(env.data.cpu().numpy(), O.clamp_cast(env_ref, np.float32))
    ref = O.logcompress(O.clamp_cast(env_ref, np.float32), 60.0, 1.0, 20000.0)
    if out_dt == "u8":
        assert np.abs(a.astype(np.int32) - O.clamp_cast(ref, np.uint8).astype(np.int32)).max() <= 1
        assert a.max() > 100
    else:
        assert np.abs(a - ref).max() <= 4e-7 * 255


def test_postprocess_rejects_bad_arguments(device):
    x = torch.zeros((10, 4), device="cuda")
    out = torch.zeros((10, 4), device="cuda")
    lib = N.lib()
    import ctypes as C
    assert lib.supra_envelope_run(C.c_void_p(x.data_ptr()), N.SUPRA_FLOAT32, 4, 10, 1, 0, C.c_void_p(out.data_ptr()),
                                  N.SUPRA_FLOAT32, None) == N.SUPRA_ERR_INVALID_ARG
    assert lib.supra_envelope_run(C.c_void_p(x.data_ptr()), N.SUPRA_FLOAT32, 4, 10, 1, 258, C.c_void_p(out.data_ptr()),
                                  N.SUPRA_FLOAT32, None) == N.SUPRA_ERR_INVALID_ARG
    assert lib.supra_envelope_run(C.c_void_p(x.data_ptr()), N.SUPRA_UINT8, 4, 10, 1, 65, C.c_void_p(out.data_ptr()),
                                  N.SUPRA_FLOAT32, None) == N.SUPRA_ERR_UNSUPPORTED
    assert lib.supra_logcompress_run(C.c_void_p(x.data_ptr()), N.SUPRA_FLOAT32, 40, 80.0, 1.0, 1024.0,
                                     C.c_void_p(out.data_ptr()), N.SUPRA_INT16, None) == N.SUPRA_ERR_UNSUPPORTED
    assert lib.supra_envelope_run(C.c_void_p(x.data_ptr()), N.SUPRA_FLOAT32, 4, 10, 70000, 65,
                                  C.c_void_p(out.data_ptr()), N.SUPRA_FLOAT32, None) == N.SUPRA_ERR_TOO_LARGE
    assert lib.supra_envelope_run(C.c_void_p(x.data_ptr()), N.SUPRA_FLOAT32, 4, 10, 0, 65,
                                  C.c_void_p(out.data_ptr()), N.SUPRA_FLOAT32, None) == N.SUPRA_OK


@pytest.mark.parametrize("n_taps,fused", [(65, False), (64, True), (17, True)])
def test_envelope_frame_batch_equals_single_frames(device, n_taps, fused):
    """A batch of frames in one launch: each frame is its own image (the FIR
    does not cross frames), bit for bit the single-frame call and the oracle."""
    from supra_amd.postprocess import HilbertFirEnvelope, envelope_log_compress
    rng = np.random.default_rng(n_taps)
    x = (rng.normal(size=(5, 203, 150)) * 3000).astype(np.float32)
    dev = torch.from_numpy(x).cuda()
    if fused:
        batch = envelope_log_compress(dev, n_taps, 70.0, 1.2, 9000.0, out_dtype=torch.float32).data
        singles = [envelope_log_compress(dev[f], n_taps, 70.0, 1.2, 9000.0, out_dtype=torch.float32).data
                   for f in range(5)]
    else:
        batch = HilbertFirEnvelope(n_taps).demodulate(dev, torch.float32).data
        singles = [HilbertFirEnvelope(n_taps).demodulate(dev[f], torch.float32).data for f in range(5)]
    torch.cuda.synchronize()
    got = batch.cpu().numpy()
    assert got.shape == x.shape
    for f in range(5):
        assert np.array_equal(got[f], singles[f].cpu().numpy())
        env = O.clamp_cast(O.envelope(x[f], n_taps), np.float32)
        if not fused:
            assert np.array_equal(got[f], env)
