"""ctypes binding of libsupra_bf.so (the C ABI declared in include/supra_bf.h
and include/supra_synth.h).

The library must be built in-tree (``python -m supra_amd.build`` or
``__graft_entry__.build()``); loading fails loudly otherwise — there is no
fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libsupra_bf.so"

# status codes / enums (supra_bf.h)
SUPRA_OK = 0
SUPRA_ERR_INVALID_ARG, SUPRA_ERR_UNSUPPORTED, SUPRA_ERR_HIP, SUPRA_ERR_NO_DEVICE, SUPRA_ERR_TOO_LARGE = -1, -2, -3, -4, -5
NOALLOC_SCRATCH_BYTES = 8 << 20     # SUPRA_NOALLOC_SCRATCH_BYTES
SUPRA_INT16 = 1
SUPRA_FLOAT32 = 2
SUPRA_UINT8 = 3
WINDOW_RECTANGULAR, WINDOW_HANN, WINDOW_HAMMING, WINDOW_GAUSS = 0, 1, 2, 3
BF_DELAY_AND_SUM, BF_DELAY_AND_STD_DEV, BF_COHERENCE_FACTOR, BF_TEST_SIGNAL = 0, 1, 2, 3
SCAN_LINEAR, SCAN_PHASED, SCAN_BIPHASED, SCAN_PLANEWAVE = 0, 1, 2, 3
TRANSDUCER_LINEAR, TRANSDUCER_PLANAR = 0, 1
WINDOW_NUM_ENTRIES = 64


class MockMetaC(C.Structure):
    """supra_mock_meta (supra_bf.h)."""
    _fields_ = [("num_elements", C.c_uint32), ("element_layout", C.c_uint32 * 2), ("num_channels", C.c_uint32),
                ("num_samples", C.c_uint32), ("num_tx", C.c_uint32), ("rx_scanline_layout", C.c_uint32 * 2),
                ("num_depths", C.c_uint32), ("depth", C.c_double), ("sampling_frequency_hz", C.c_double),
                ("speed_of_sound_mm_per_s", C.c_double)]


class SupraError(RuntimeError):
    """A SUPRA_ERR_* status returned by the library, with its message."""

    def __init__(self, code: int, message: str):
        super().__init__(f"supra_bf error {code}: {message}")
        self.code = code


class TxParams(C.Structure):
    _fields_ = [
        ("first_active_element", C.c_uint16 * 2),
        ("last_active_element", C.c_uint16 * 2),
        ("tx_scanline_idx", C.c_uint16),
        ("reserved", C.c_uint16 * 3),
        ("initial_delay", C.c_double),
    ]


class ScanlineRx(C.Structure):
    _fields_ = [
        ("position", C.c_double * 3),
        ("direction", C.c_double * 3),
        ("tx_weights", C.c_double * 4),
        ("tx", TxParams * 4),
        ("max_element_distance", C.c_double * 2),
    ]


class RxGeometryC(C.Structure):
    _fields_ = [
        ("num_rx_scanlines", C.c_uint32),
        ("rx_scanline_layout", C.c_uint32 * 2),
        ("num_depths", C.c_uint32),
        ("num_elements", C.c_uint32),
        ("element_layout", C.c_uint32 * 2),
        ("scanlines", C.POINTER(ScanlineRx)),
        ("depths", C.POINTER(C.c_float)),
        ("element_x", C.POINTER(C.c_float)),
        ("element_y", C.POINTER(C.c_float)),
        ("element_to_channel", C.POINTER(C.c_int32)),
        ("speed_of_sound_mm_per_s", C.c_double),
        ("depth", C.c_double),
    ]


class BeamformerConfigC(C.Structure):
    _fields_ = [
        ("scan_type", C.c_int32),
        ("transducer_type", C.c_int32),
        ("element_layout", C.c_uint32 * 2),
        ("pitch", C.c_double * 2),
        ("num_scanlines", C.c_uint32 * 2),
        ("rx_scanline_subdivision", C.c_uint32 * 2),
        ("max_aperture", C.c_uint32 * 2),
        ("tx_max_aperture", C.c_uint32 * 2),
        ("depth", C.c_double),
        ("num_depths", C.c_uint32),
        ("steering_angle_deg", C.c_double * 2),
        ("sector_angle_deg", C.c_double * 2),
        ("tx_focus_active", C.c_int32),
        ("tx_focus_depth", C.c_double),
        ("tx_focus_width", C.c_double),
        ("speed_of_sound", C.c_double),
    ]


class SynthParamsC(C.Structure):
    _fields_ = [
        ("num_tx", C.c_uint32),
        ("num_channels", C.c_uint32),
        ("num_samples", C.c_uint32),
        ("num_elements", C.c_uint32),
        ("sampling_frequency_hz", C.c_double),
        ("speed_of_sound_m_per_s", C.c_double),
        ("center_frequency_hz", C.c_double),
        ("fractional_bandwidth", C.c_double),
        ("snr_db", C.c_double),
        ("peak", C.c_double),
        ("beam_sigma_mm", C.c_double),
        ("noise_seed", C.c_uint64),
        ("num_scatterers", C.c_uint32),
        ("scatterers", C.POINTER(C.c_double)),
        ("tx_origin", C.POINTER(C.c_double)),
        ("tx_direction", C.POINTER(C.c_double)),
        ("tx_t0", C.POINTER(C.c_double)),
        ("tx_first_element", C.POINTER(C.c_uint32)),
        ("tx_last_element", C.POINTER(C.c_uint32)),
        ("element_pos", C.POINTER(C.c_double)),
    ]


# name -> (restype, argtypes); every symbol the public headers declare.
_SIGNATURES = {
    "supra_geometry_sizes": (C.c_int, [C.POINTER(BeamformerConfigC), C.POINTER(C.c_uint32),
                                       C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "supra_geometry_build": (C.c_int, [C.POINTER(BeamformerConfigC), C.POINTER(ScanlineRx),
                                       C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "supra_bf_create": (C.c_int, [C.POINTER(RxGeometryC), C.POINTER(C.c_void_p)]),
    "supra_bf_destroy": (None, [C.c_void_p]),
    "supra_das_run": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_uint32, C.c_uint32,
                                C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_int32, C.c_float,
                                C.c_int32, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p]),
    "supra_das_run_frames": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32, C.c_void_p, C.c_int32, C.c_uint32,
                                       C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_int32,
                                       C.c_float, C.c_int32, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p]),
    "supra_das_forward_c": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_double, C.c_double, C.c_void_p, C.c_int32, C.c_float, C.c_int32,
                                      C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "supra_das_forward_c_noalloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32,
                                              C.c_uint32, C.c_double, C.c_double, C.c_void_p, C.c_int32, C.c_float,
                                              C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "supra_das_adjoint": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.c_double, C.c_double, C.c_void_p, C.c_int32, C.c_float, C.c_int32,
                                    C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "supra_rawdelay_run": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_uint32, C.c_uint32, C.c_uint32,
                                     C.c_double, C.c_double, C.c_double, C.c_int32, C.c_float, C.c_void_p,
                                     C.c_int32, C.POINTER(C.c_double), C.c_void_p]),
    "supra_mv_run": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                               C.c_double, C.c_int32, C.c_double, C.c_void_p, C.c_int32, C.c_void_p]),
    "supra_mvpcg_run": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                  C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_int32, C.c_void_p, C.c_int32,
                                  C.c_void_p]),
    "supra_mv_das_run": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_uint32, C.c_uint32, C.c_uint32,
                                   C.c_double, C.c_double, C.c_double, C.c_int32, C.c_float, C.c_uint32,
                                   C.c_uint32, C.c_double, C.c_int32, C.c_double, C.c_void_p, C.c_int32,
                                   C.c_void_p]),
    "supra_mock_read_meta": (C.c_int, [C.c_char_p, C.c_void_p]),
    "supra_mock_read_geometry": (C.c_int, [C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "supra_mock_write_json": (C.c_int, [C.c_char_p, C.c_void_p, C.POINTER(RxGeometryC)]),
    "supra_mock_write_ascii": (C.c_int, [C.c_char_p, C.c_void_p, C.POINTER(RxGeometryC)]),
    "supra_mock_count_frames": (C.c_int, [C.c_char_p, C.c_void_p, C.POINTER(C.c_uint64)]),
    "supra_mock_read_frames": (C.c_int, [C.c_char_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p]),
    "supra_envelope_run": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                     C.c_void_p,
                                     C.c_int32, C.c_void_p]),
    "supra_logcompress_run": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint64, C.c_double, C.c_double, C.c_double,
                                        C.c_void_p, C.c_int32, C.c_void_p]),
    "supra_envelope_logcompress_run": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32, C.c_uint32, C.c_uint32,
                                                 C.c_uint32,
                                                 C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_int32,
                                                 C.c_void_p]),
    "supra_scanconv_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.c_double, C.c_uint32, C.c_double,
                                        C.c_uint32, C.POINTER(C.c_void_p)]),
    "supra_scanconv_destroy": (None, [C.c_void_p]),
    "supra_scanconv_geometry": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_double),
                                          C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "supra_scanconv_tables": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "supra_scanconv_mask": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "supra_scanconv_run": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_uint32, C.c_void_p, C.c_int32,
                                     C.c_void_p]),
    "supra_batch_create": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int32, C.POINTER(C.c_void_p)]),
    "supra_batch_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "supra_batch_latencies": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "supra_batch_config": (C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t]),
    "supra_batch_destroy": (None, [C.c_void_p]),
    "supra_window_lut": (C.c_int, [C.c_int32, C.c_float, C.POINTER(C.c_float)]),
    "supra_last_error": (C.c_char_p, []),
    "supra_version": (C.c_char_p, []),
    "supra_synth_uniform": (C.c_double, [C.c_uint64, C.c_uint64]),
    "supra_synth_rf_host": (C.c_int, [C.POINTER(SynthParamsC), C.c_void_p, C.c_int32]),
    "supra_synth_rf_device": (C.c_int, [C.POINTER(SynthParamsC), C.c_void_p, C.c_int32, C.c_void_p]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libsupra_bf.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m supra_amd.build` "
                "(the beamformer has no CPU fallback)")
        handle = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | C.RTLD_GLOBAL)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


def check(status: int) -> None:
    if status != SUPRA_OK:
        msg = lib().supra_last_error()
        raise SupraError(status, msg.decode() if msg else "")


def ptr(array, ctype):
    """ctypes pointer to a contiguous numpy array (or None)."""
    if array is None:
        return None
    return array.ctypes.data_as(C.POINTER(ctype))
