"""Batched beamforming of many acquisitions of uneven size (supra_batch_*):
one graph launch runs every acquisition's delay-and-sum or pre-delay +
minimum variance on persistent work-queue kernels and reports per-acquisition
completion times measured on the device.  No reference counterpart: SUPRA
beamforms one USRawData per node call (BeamformingNode.cpp:128-177)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .beamformer import RxBeamformer, _dtype_code, _stream_handle

DAS = 0
MV = 1


class BatchItemC(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("rf", C.c_void_p), ("rf_dtype", C.c_int32), ("mode", C.c_int32),
                ("num_tx", C.c_uint32), ("num_channels", C.c_uint32), ("num_samples", C.c_uint32),
                ("window_type", C.c_int32), ("sampling_frequency_hz", C.c_double), ("f_number", C.c_double),
                ("speed_of_sound_mm_per_s", C.c_double), ("window_parameter", C.c_float),
                ("interpolate_transmits", C.c_int32), ("additional_offset", C.c_int32),
                ("sub_array_size", C.c_uint32), ("temporal_smoothing", C.c_uint32), ("compute_means", C.c_int32),
                ("sub_array_scaling_power", C.c_double), ("loading_factor", C.c_double), ("out", C.c_void_p),
                ("out_dtype", C.c_int32), ("reserved", C.c_int32)]


class Batch:
    """A planned, graph-captured batch.  Each entry of `items` is a dict with
    engine (RxBeamformer of the acquisition's geometry), rf (device tensor
    [tx][ch][s]), out (device tensor [depth][scanline]), fs_hz, mode (DAS/MV)
    and optional f_number, speed_of_sound_mm_per_s, window_type,
    window_parameter, interpolate_transmits, additional_offset,
    sub_array_size, temporal_smoothing, compute_means,
    sub_array_scaling_power, loading_factor."""

    def __init__(self, items: list, autotune: bool = True):
        arr = (BatchItemC * len(items))()
        self._keep = []
        for c, it in zip(arr, items):
            rf, out, eng = it["rf"], it["out"], it["engine"]
            if not (rf.is_cuda and rf.is_contiguous() and rf.dim() == 3):
                raise ValueError("rf must be a contiguous device tensor [tx][ch][s]")
            if not (out.is_cuda and out.is_contiguous()):
                raise ValueError("out must be a contiguous device tensor")
            g = eng.geometry
            if tuple(out.shape) != (g.num_depths, g.num_rx):
                # the kernels write num_depths x num_rx samples through the bare pointer
                raise ValueError(f"out must have shape {(g.num_depths, g.num_rx)} (the engine's image), "
                                 f"got {tuple(out.shape)}")
            c.ctx = eng._ctx.value
            c.rf = rf.data_ptr()
            c.rf_dtype = _dtype_code(rf.dtype)
            c.mode = it.get("mode", DAS)
            c.num_tx, c.num_channels, c.num_samples = (int(v) for v in rf.shape)
            c.window_type = it.get("window_type", N.WINDOW_HANN)
            c.sampling_frequency_hz = it["fs_hz"]
            c.f_number = it.get("f_number", 1.0)
            c.speed_of_sound_mm_per_s = it.get("speed_of_sound_mm_per_s", 1540e3)
            c.window_parameter = it.get("window_parameter", 0.0)
            c.interpolate_transmits = int(bool(it.get("interpolate_transmits", False)))
            c.additional_offset = it.get("additional_offset", 0)
            c.sub_array_size = it.get("sub_array_size", 0)
            c.temporal_smoothing = it.get("temporal_smoothing", 3)
            c.compute_means = int(bool(it.get("compute_means", False)))
            c.sub_array_scaling_power = it.get("sub_array_scaling_power", 1.5)
            c.loading_factor = it.get("loading_factor", 1.0)
            c.out = out.data_ptr()
            c.out_dtype = _dtype_code(out.dtype)
            self._keep += [rf, out, eng]
        self.n = len(items)
        h = C.c_void_p()
        N.check(N.lib().supra_batch_create(arr, len(items), int(bool(autotune)), C.byref(h)))
        self._h = h

    def launch(self, stream=None) -> None:
        N.check(N.lib().supra_batch_launch(self._h, C.c_void_p(_stream_handle(stream))))

    def latencies(self):
        """(per-acquisition completion time after launch start [us], the largest)."""
        out = np.zeros(self.n)
        total = C.c_double()
        N.check(N.lib().supra_batch_latencies(self._h, out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(total)))
        return out, total.value

    def config(self) -> str:
        buf = C.create_string_buffer(4096)
        N.check(N.lib().supra_batch_config(self._h, buf, len(buf)))
        return buf.value.decode()

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            N.lib().supra_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
