"""After the beamformer: envelope detection and log compression (the B-mode
stages that consume a ``USImage``), mirroring SUPRA's engines and nodes:

* ``HilbertFirEnvelope`` — HilbertFirEnvelope (HilbertFirEnvelope.h,
  HilbertFirEnvelope.cu:63-104): ``demodulate(image)``;
* ``LogCompressor`` — LogCompressor (LogCompressor.h,
  LogCompressor.cu:43-64): ``compress(image, dynamic_range, scale, in_max)``;
* ``HilbertFirEnvelopeNode`` / ``LogCompressorNode`` — the nodes' XML
  parameters with the reference's names, ranges and defaults
  (HilbertFirEnvelopeNode.cpp:42-43, LogCompressorNode.cpp:42-45);
* ``envelope_log_compress`` — both in one launch (no intermediate image).

All computation is in libsupra_bf.so (supra_envelope_run,
supra_logcompress_run, supra_envelope_logcompress_run).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _native as N
from .beamformer import USImage, _Param, _stream_handle

_CODES = {torch.int16: N.SUPRA_INT16, torch.float32: N.SUPRA_FLOAT32, torch.uint8: N.SUPRA_UINT8}
ENVELOPE_OUTPUT_TYPES = {"float": torch.float32, "int16": torch.int16}
LOG_OUTPUT_TYPES = {"float": torch.float32, "uint8": torch.uint8}


def _code(dtype) -> int:
    if dtype not in _CODES:
        raise TypeError(f"unsupported element type {dtype}")
    return _CODES[dtype]


def _image(img):
    """Device tensor [sample][scanline] or a batch [frame][sample][scanline]."""
    data = img.data if isinstance(img, USImage) else img
    if data.dim() not in (2, 3) or not data.is_cuda or not data.is_contiguous():
        raise ValueError("image must be a contiguous device tensor [sample][scanline] "
                         "or [frame][sample][scanline]")
    return data


def _dims(data):
    """(frames, samples, scanlines) of an image or batch."""
    return (1, *(int(v) for v in data.shape)) if data.dim() == 2 else tuple(int(v) for v in data.shape)


class HilbertFirEnvelope:
    """Envelope sqrt(x^2 + H{x}^2) along each scanline, H the Hamming-windowed
    FIR Hilbert transformer of `filter_length` taps."""

    def __init__(self, filter_length: int = 65):
        self.filter_length = int(filter_length)

    def demodulate(self, image, out_dtype=torch.float32, out=None, stream=None) -> USImage:
        data = _image(image)
        n_f, n_s, n_rx = _dims(data)
        if out is None:
            out = torch.empty(tuple(data.shape), dtype=out_dtype, device=data.device)
        elif tuple(out.shape) != tuple(data.shape) or out.dtype != out_dtype or not out.is_contiguous():
            raise ValueError(f"out must be a contiguous {out_dtype} tensor of shape {tuple(data.shape)}")
        N.check(N.lib().supra_envelope_run(C.c_void_p(data.data_ptr()), _code(data.dtype), n_rx, n_s, n_f,
                                           self.filter_length, C.c_void_p(out.data_ptr()), _code(out_dtype),
                                           C.c_void_p(_stream_handle(stream))))
        props = dict(image.properties) if isinstance(image, USImage) else {}
        props.update({"imageState": "EnvDetected", "HilbertFirEnvelope.FilterLength": self.filter_length})
        return USImage(out, (n_rx, n_s), props)


class LogCompressor:
    """log10(|x| 10^(dB/20)/inMax + 1) * scale * 255 / log10(10^(dB/20) + 1)."""

    def compress(self, image, dynamic_range: float, scale: float, in_max: float, out_dtype=torch.float32,
                 out=None, stream=None) -> USImage:
        data = image.data if isinstance(image, USImage) else image
        if not data.is_cuda or not data.is_contiguous():
            raise ValueError("image must be a contiguous device tensor")
        if out is None:
            out = torch.empty(tuple(data.shape), dtype=out_dtype, device=data.device)
        elif tuple(out.shape) != tuple(data.shape) or out.dtype != out_dtype or not out.is_contiguous():
            raise ValueError(f"out must be a contiguous {out_dtype} tensor of shape {tuple(data.shape)}")
        N.check(N.lib().supra_logcompress_run(C.c_void_p(data.data_ptr()), _code(data.dtype), data.numel(),
                                              float(dynamic_range), float(scale), float(in_max),
                                              C.c_void_p(out.data_ptr()), _code(out_dtype),
                                              C.c_void_p(_stream_handle(stream))))
        size = image.size if isinstance(image, USImage) else tuple(reversed(tuple(data.shape)))
        props = dict(image.properties) if isinstance(image, USImage) else {}
        props.update({"imageState": "Compressed", "LogCompressor.DynamicRange": dynamic_range,
                      "LogCompressor.Gain": scale, "LogCompressor.InMax": in_max})
        return USImage(out, size, props)


def envelope_log_compress(image, filter_length: int = 65, dynamic_range: float = 80.0, scale: float = 1.0,
                          in_max: float = 1024.0, out_dtype=torch.uint8, out=None, stream=None) -> USImage:
    """HilbertFirEnvelope (float envelope) then LogCompressor in one launch."""
    data = _image(image)
    n_f, n_s, n_rx = _dims(data)
    if out is None:
        out = torch.empty(tuple(data.shape), dtype=out_dtype, device=data.device)
    elif tuple(out.shape) != tuple(data.shape) or out.dtype != out_dtype or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous {out_dtype} tensor of shape {tuple(data.shape)}")
    N.check(N.lib().supra_envelope_logcompress_run(C.c_void_p(data.data_ptr()), _code(data.dtype), n_rx, n_s, n_f,
                                                   int(filter_length), float(dynamic_range), float(scale),
                                                   float(in_max), C.c_void_p(out.data_ptr()), _code(out_dtype),
                                                   C.c_void_p(_stream_handle(stream))))
    props = dict(image.properties) if isinstance(image, USImage) else {}
    props.update({"imageState": "Compressed"})
    return USImage(out, (n_rx, n_s), props)


class _Node:
    PARAMETERS: dict = {}

    def __init__(self, node_id: str, **params):
        self.node_id = node_id
        self.config = {k: p.default for k, p in self.PARAMETERS.items()}
        for k, v in params.items():
            self.set(k, v)

    def set(self, key: str, value) -> None:
        if key not in self.PARAMETERS:
            raise KeyError(f"{type(self).__name__} has no parameter {key!r}")
        self.config[key] = self.PARAMETERS[key].validate(key, value)


class HilbertFirEnvelopeNode(_Node):
    """HilbertFirEnvelopeNode (HilbertFirEnvelopeNode.cpp:24-160)."""

    PARAMETERS = {
        "filterLength": _Param(65, 1, 257),
        "outputType": _Param("float", choices=set(ENVELOPE_OUTPUT_TYPES)),
    }

    def __init__(self, node_id: str = "HilbertFirEnvelopeNode", **params):
        super().__init__(node_id, **params)

    def process(self, image: USImage, stream=None) -> USImage:
        env = HilbertFirEnvelope(self.config["filterLength"])
        return env.demodulate(image, ENVELOPE_OUTPUT_TYPES[self.config["outputType"]], stream=stream)


class LogCompressorNode(_Node):
    """LogCompressorNode (LogCompressorNode.cpp:24-160)."""

    PARAMETERS = {
        "dynamicRange": _Param(80.0, 1.0, 200.0),
        "gain": _Param(1.0, 0.01, 2.0),
        "inMax": _Param(1024.0, 10.0, 50000.0),
        "outputType": _Param("float", choices=set(LOG_OUTPUT_TYPES)),
    }

    def __init__(self, node_id: str = "LogCompressorNode", **params):
        super().__init__(node_id, **params)

    def process(self, image: USImage, stream=None) -> USImage:
        return LogCompressor().compress(image, self.config["dynamicRange"], self.config["gain"],
                                        self.config["inMax"], LOG_OUTPUT_TYPES[self.config["outputType"]],
                                        stream=stream)
