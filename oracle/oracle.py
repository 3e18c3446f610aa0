"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the supra_amd package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        h = C.CDLL(str(LIB_PATH))
        h.oracle_window_lut.restype = C.c_int
        h.oracle_window_lut.argtypes = [C.c_int, C.c_float, C.c_int, C.c_void_p]
        h.oracle_das.restype = C.c_int
        h.oracle_das.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32,
                                 C.c_double, C.c_double, C.c_double, C.c_int, C.c_float, C.c_int, C.c_int32,
                                 C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_uint64, C.c_void_p, C.c_int]
        h.oracle_rawdelay.restype = C.c_int
        h.oracle_rawdelay.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_double, C.c_double, C.c_double, C.c_int, C.c_float, C.c_int, C.c_int,
                                      C.c_void_p, C.c_int]
        h.oracle_das_adjoint.restype = C.c_int
        h.oracle_das_adjoint.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                         C.c_double, C.c_double, C.c_void_p, C.c_int, C.c_float, C.c_int, C.c_int32,
                                         C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        h.oracle_mv.restype = C.c_int
        h.oracle_mv.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                C.c_double, C.c_int, C.c_double, C.c_int, C.c_void_p, C.c_int, C.c_uint32,
                                C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
        h.oracle_mvpcg.restype = C.c_int
        h.oracle_mvpcg.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                   C.c_uint32, C.c_double, C.c_double, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                   C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
        h.oracle_envelope.restype = C.c_int
        h.oracle_envelope.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_void_p,
                                      C.c_int]
        h.oracle_hilbert_filter.restype = C.c_int
        h.oracle_hilbert_filter.argtypes = [C.c_uint32, C.c_int, C.c_void_p]
        h.oracle_logcompress.restype = C.c_int
        h.oracle_logcompress.argtypes = [C.c_void_p, C.c_int, C.c_uint64, C.c_double, C.c_double, C.c_double,
                                         C.c_int, C.c_void_p, C.c_int]
        h.oracle_das_decision_diff.restype = C.c_int
        h.oracle_das_decision_diff.argtypes = [C.c_void_p, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_int,
                                               C.c_double, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int32,
                                               C.c_void_p, C.c_int]
        _lib = h
    return _lib


def default_threads() -> int:
    """Every core this process may run on."""
    return max(1, len(os.sched_getaffinity(0)))


def window_lut(window_type: int, window_parameter: float, precision: int = 32) -> np.ndarray:
    out = np.zeros(64, dtype=np.float32 if precision == 32 else np.float64)
    assert lib().oracle_window_lut(window_type, window_parameter, precision, out.ctypes.data) == 0
    return out


def das(geom, rf: np.ndarray, fs_hz: float, f_number: float, speed_of_sound_mm_per_s: float,
        window_type: int, window_parameter: float = 0.0, interpolate_transmits: bool = False,
        additional_offset: int = 0, test_signal: bool = False, out_int16: bool = False,
        precision: int = 32, pixels: np.ndarray | None = None, threads: int | None = None,
        sample_beamformer: int = 0) -> np.ndarray:
    """Oracle DAS image [num_depths, num_rx] (or the values at `pixels`,
    flat indices depth * num_rx + scanline).  `geom` is a RxGeometry.
    sample_beamformer: 0 DelayAndSum, 1 DelayAndStdDev, 2
    CoherenceFactorDelayAndSum (test_signal=True selects 3, TestSignal)."""
    rf = np.ascontiguousarray(rf)
    assert rf.ndim == 3 and rf.dtype in (np.int16, np.float32)
    gc = geom.to_c()
    dt = np.float32 if precision == 32 else np.float64
    if pixels is None:
        out = np.zeros((geom.num_depths, geom.num_rx), dtype=dt)
        pix_ptr, npix = None, 0
    else:
        pixels = np.ascontiguousarray(pixels, dtype=np.uint32)
        out = np.zeros(pixels.shape[0], dtype=dt)
        pix_ptr, npix = pixels.ctypes.data, pixels.shape[0]
    code = lib().oracle_das(C.addressof(gc), rf.ctypes.data, 1 if rf.dtype == np.int16 else 2,
                            rf.shape[0], rf.shape[1], rf.shape[2], fs_hz, f_number, speed_of_sound_mm_per_s,
                            window_type, window_parameter, int(interpolate_transmits), additional_offset,
                            3 if test_signal else int(sample_beamformer), 1 if out_int16 else 2, precision,
                            pix_ptr, npix,
                            out.ctypes.data, threads or default_threads())
    if code != 0:
        raise ValueError(f"oracle_das rejected its arguments ({code})")
    return out


def das_adjoint(geom, rf: np.ndarray, fs_hz: float, f_number: float, c_mm_per_s, window_type: int,
                grad_image: np.ndarray, window_parameter: float = 0.0, interpolate_transmits: bool = False,
                additional_offset: int = 0, precision: int = 64, want_rf: bool = True, want_norm: bool = False,
                threads: int | None = None):
    """Oracle reverse mode of the 2D DAS: rf float32 [frame][tx][ch][s], one
    speed of sound per frame, grad_image [frame][depth][scanline].
    Returns (grad_rf or None, grad_c [frame] in 1/(mm/s), norm or None)."""
    rf = np.ascontiguousarray(rf, dtype=np.float32)
    if rf.ndim == 3:
        rf = rf[None]
    n_f = rf.shape[0]
    c = np.ascontiguousarray(np.broadcast_to(np.asarray(c_mm_per_s, dtype=np.float64), (n_f,)))
    gi = np.ascontiguousarray(np.broadcast_to(np.asarray(grad_image, dtype=np.float64),
                                              (n_f, geom.num_depths, geom.num_rx)))
    gc = geom.to_c()
    dt = np.float32 if precision == 32 else np.float64    # 6432: FP64 values on FP32 decisions
    grad_rf = np.zeros(rf.shape, dtype=dt) if want_rf else None
    grad_c = np.zeros(n_f)
    slots = 4 if interpolate_transmits else 1
    norm = np.zeros((n_f, slots, geom.num_rx, geom.num_depths), dtype=dt) if want_norm else None
    code = lib().oracle_das_adjoint(C.addressof(gc), rf.ctypes.data, n_f, rf.shape[1], rf.shape[2], rf.shape[3],
                                    fs_hz, f_number, c.ctypes.data, window_type, window_parameter,
                                    int(interpolate_transmits), additional_offset, gi.ctypes.data, precision,
                                    grad_rf.ctypes.data if want_rf else None, grad_c.ctypes.data,
                                    norm.ctypes.data if want_norm else None, threads or default_threads())
    if code != 0:
        raise ValueError(f"oracle_das_adjoint rejected its arguments ({code})")
    return grad_rf, grad_c, norm


def das_decision_diff(geom, num_samples: int, fs_hz: float, f_number: float, speed_of_sound_mm_per_s: float,
                      window_type: int, window_parameter: float = 0.0, interpolate_transmits: bool = False,
                      additional_offset: int = 0, threads: int | None = None, precision_a: int = 32,
                      precision_b: int = 64, speed_of_sound_b_mm_per_s: float | None = None) -> np.ndarray:
    """uint8 [num_depths, num_rx]: where two evaluations of the 2D DAS take
    different decisions for some element (1 membership, 2 window value, 4
    floor of the delay, 8 range test).  By default FP32 against FP64 at one
    speed of sound; speed_of_sound_b_mm_per_s / precision_b give evaluation B."""
    gc = geom.to_c()
    out = np.zeros((geom.num_depths, geom.num_rx), dtype=np.uint8)
    c_b = speed_of_sound_mm_per_s if speed_of_sound_b_mm_per_s is None else speed_of_sound_b_mm_per_s
    code = lib().oracle_das_decision_diff(C.addressof(gc), num_samples, fs_hz, f_number, speed_of_sound_mm_per_s,
                                          precision_a, c_b, precision_b, window_type, window_parameter,
                                          int(interpolate_transmits), additional_offset, out.ctypes.data,
                                          threads or default_threads())
    if code != 0:
        raise ValueError(f"oracle_das_decision_diff rejected its arguments ({code})")
    return out


def rawdelay(geom, rf: np.ndarray, fs_hz: float, f_number: float, speed_of_sound_mm_per_s: float,
             window_type: int, window_parameter: float = 0.0, out_int16: bool = False, precision: int = 32,
             threads: int | None = None) -> np.ndarray:
    """Oracle pre-delay [num_rx, num_channels, num_depths] (2D and 3D geometries)."""
    rf = np.ascontiguousarray(rf)
    gc = geom.to_c()
    out = np.zeros((geom.num_rx, rf.shape[1], geom.num_depths), dtype=np.float32 if precision == 32 else np.float64)
    code = lib().oracle_rawdelay(C.addressof(gc), rf.ctypes.data, 1 if rf.dtype == np.int16 else 2, rf.shape[0],
                                 rf.shape[1], rf.shape[2], fs_hz, f_number, speed_of_sound_mm_per_s, window_type,
                                 window_parameter, 1 if out_int16 else 2, precision, out.ctypes.data,
                                 threads or default_threads())
    if code != 0:
        raise ValueError(f"oracle_rawdelay rejected its arguments ({code})")
    return out


def mv(x: np.ndarray, sub_array_size: int = 0, temporal_smoothing: int = 3, scaling_power: float = 1.5,
       compute_means: bool = False, loading_factor: float = 1.0, out_int16: bool = False,
       threads: int | None = None, debug: tuple | None = None):
    """Oracle minimum variance (FP64) on x [scanline, channel, sample] -> [sample, scanline].
    debug=(scanline, sample) also returns (A, w, (first, last, L_t, K_act)) of that sample."""
    x = np.ascontiguousarray(x)
    assert x.ndim == 3 and x.dtype in (np.int16, np.float32)
    n_rx, n_ch, n_d = x.shape
    L = sub_array_size or n_ch // 2
    out = np.zeros((n_d, n_rx))
    dA = np.zeros(L * L)
    dw = np.zeros(L)
    dm = np.zeros(5, dtype=np.int32)
    ds, dt = debug if debug else (0xFFFFFFFF, 0xFFFFFFFF)
    code = lib().oracle_mv(x.ctypes.data, 1 if x.dtype == np.int16 else 2, n_rx, n_ch, n_d, sub_array_size,
                           temporal_smoothing, scaling_power, int(compute_means), loading_factor,
                           1 if out_int16 else 2, out.ctypes.data, threads or default_threads(), ds, dt,
                           dA.ctypes.data, dw.ctypes.data, dm.ctypes.data)
    if code != 0:
        raise ValueError(f"oracle_mv rejected its arguments ({code})")
    if debug:
        n = int(dm[2])
        return out, dA[:n * n].reshape(n, n), dw[:n], tuple(int(v) for v in dm[:4])
    return out


def mvpcg(x: np.ndarray, sub_array_size: int = 0, temporal_smoothing: int = 3, max_iterations: int = 0,
          threshold: float = 1e-4, scaling_power: float = 1.5, compute_means: bool = False, out_int16: bool = False,
          threads: int | None = None, debug: tuple | None = None):
    """Oracle minimum variance with RxBeamformerMVpcg's PCG solve (FP64).
    debug=(scanline, sample) also returns (A, w, (first, last, L_t, K_act,
    iterations)) of that sample."""
    x = np.ascontiguousarray(x)
    assert x.ndim == 3 and x.dtype in (np.int16, np.float32)
    n_rx, n_ch, n_d = x.shape
    L = sub_array_size or n_ch // 2
    out = np.zeros((n_d, n_rx))
    dA = np.zeros(L * L)
    dw = np.zeros(L)
    dm = np.zeros(5, dtype=np.int32)
    ds, dt = debug if debug else (0xFFFFFFFF, 0xFFFFFFFF)
    code = lib().oracle_mvpcg(x.ctypes.data, 1 if x.dtype == np.int16 else 2, n_rx, n_ch, n_d, sub_array_size,
                              temporal_smoothing, max_iterations, threshold, scaling_power, int(compute_means),
                              1 if out_int16 else 2, out.ctypes.data, threads or default_threads(), ds, dt,
                              dA.ctypes.data, dw.ctypes.data, dm.ctypes.data)
    if code != 0:
        raise ValueError(f"oracle_mvpcg rejected its arguments ({code})")
    if debug:
        n = int(dm[2])
        return out, dA[:n * n].reshape(n, n), dw[:n], tuple(int(v) for v in dm)
    return out


_DTYPE = {np.dtype(np.int16): 1, np.dtype(np.float32): 2, np.dtype(np.uint8): 3}


def hilbert_filter(filter_length: int, precision: int = 32) -> np.ndarray:
    """Hilbert transformer FIR taps with the Hamming window (FirFilterFactory)."""
    out = np.zeros(filter_length, dtype=np.float32 if precision == 32 else np.float64)
    assert lib().oracle_hilbert_filter(filter_length, precision, out.ctypes.data) == 0
    return out


def envelope(image: np.ndarray, filter_length: int = 65, precision: int = 32, threads: int | None = None):
    """Envelope sqrt(x^2 + H{x}^2) of an image [sample][scanline] (before the
    output cast), HilbertFirEnvelope's FIR along the samples."""
    image = np.ascontiguousarray(image)
    n_s, n_rx = image.shape
    out = np.zeros(image.shape, dtype=np.float32 if precision == 32 else np.float64)
    code = lib().oracle_envelope(image.ctypes.data, _DTYPE[image.dtype], n_rx, n_s, filter_length, precision,
                                 out.ctypes.data, threads or default_threads())
    if code != 0:
        raise ValueError(f"oracle_envelope rejected its arguments ({code})")
    return out


def logcompress(values: np.ndarray, dynamic_range_db: float = 80.0, gain: float = 1.0, in_max: float = 1024.0,
                precision: int = 32, threads: int | None = None):
    """LogCompressor's log10(|x| inScale + 1) * gain * 255 / log10(DR + 1)
    (before the output cast), same shape as values."""
    values = np.ascontiguousarray(values)
    out = np.zeros(values.shape, dtype=np.float32 if precision == 32 else np.float64)
    code = lib().oracle_logcompress(values.ctypes.data, _DTYPE[values.dtype], values.size, dynamic_range_db, gain,
                                    in_max, precision, out.ctypes.data, threads or default_threads())
    if code != 0:
        raise ValueError(f"oracle_logcompress rejected its arguments ({code})")
    return out


def clamp_cast(values: np.ndarray, dtype) -> np.ndarray:
    """clampCast (utilities/cudaUtility.h:111-147): clamp to the type's
    limits (int16: +-32767, uint8: 0..255, float: +-FLT_MAX; NaN -> the lower
    limit) and truncate."""
    v = np.asarray(values)
    lo, hi = {np.dtype(np.int16): (-32767.0, 32767.0), np.dtype(np.uint8): (0.0, 255.0),
              np.dtype(np.float32): (-3.4028234663852886e38, 3.4028234663852886e38)}[np.dtype(dtype)]
    c = np.fmin(np.fmax(v, v.dtype.type(lo)), v.dtype.type(hi))
    c = np.where(np.isnan(v), lo, c)
    return c.astype(dtype)
