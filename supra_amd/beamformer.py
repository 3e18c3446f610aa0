"""Host-side mirror of SUPRA's beamforming engine and node interfaces.

* ``USRawData`` / ``USImage`` — the records that flow into and out of the
  node (src/SupraLib/Beamformer/USRawData.h:26-81, src/SupraLib/USImage.h).
  Payloads are torch tensors in device memory (torch is only the allocator
  and stream provider here; all computation is in libsupra_bf.so).
* ``RxBeamformer`` — RxBeamformerCuda (RxBeamformerCuda.h:28-80): built from
  the receive geometry, ``perform_rx_beamforming`` has the reference's
  argument list and calls supra_das_run.
* ``BeamformingNode`` — BeamformingNode (BeamformingNode.cpp:24-232): the
  XML parameters with the reference's names, ranges and defaults, and the
  "new engine when the geometry object changes" rule.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .geometry import RxGeometry

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None

WINDOW_TYPES = {"Rectangular": N.WINDOW_RECTANGULAR, "Hann": N.WINDOW_HANN,
                "Hamming": N.WINDOW_HAMMING, "Gauss": N.WINDOW_GAUSS}
BEAMFORMER_TYPES = {"DelayAndSum": N.BF_DELAY_AND_SUM, "DelayAndStdDev": N.BF_DELAY_AND_STD_DEV,
                    "CoherenceFactorDelayAndSum": N.BF_COHERENCE_FACTOR, "TestSignal": N.BF_TEST_SIGNAL}
OUTPUT_TYPES = {"float": N.SUPRA_FLOAT32, "int16": N.SUPRA_INT16}


def _dtype_code(t) -> int:
    if t == torch.int16:
        return N.SUPRA_INT16
    if t == torch.float32:
        return N.SUPRA_FLOAT32
    raise TypeError(f"unsupported element type {t} (int16 or float32)")


def _stream_handle(stream) -> int:
    if stream is None:
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


@dataclass
class USRawData:
    """RF record: data [num_tx][num_channels][num_samples] on the GPU."""
    data: "torch.Tensor"
    sampling_frequency: float                 # Hz
    geometry: RxGeometry
    num_elements: int = 0
    element_layout: tuple = (0, 0)

    def __post_init__(self):
        if self.num_elements == 0:
            self.num_elements = self.geometry.num_elements
            self.element_layout = self.geometry.element_layout

    @property
    def num_scanlines(self) -> int:
        return int(self.data.shape[-3])

    @property
    def num_received_channels(self) -> int:
        return int(self.data.shape[-2])

    @property
    def num_samples(self) -> int:
        return int(self.data.shape[-1])


@dataclass
class USImage:
    """Beamformed image: data [num_depths][num_scanlines] (scanline fastest).
    properties["geometry"], when present, is the RxGeometry whose scanlines
    the image holds (USImageProperties::getScanlineInfo)."""
    data: "torch.Tensor"
    size: tuple                               # (num_scanlines, num_depths) as in USImage
    properties: dict = field(default_factory=dict)


class RxBeamformer:
    """RxBeamformerCuda equivalent on gfx950 (owns a supra_bf_ctx)."""

    DelayAndSum = N.BF_DELAY_AND_SUM
    TestSignal = N.BF_TEST_SIGNAL

    def __init__(self, geometry: RxGeometry):
        self.geometry = geometry
        self._gc = geometry.to_c()
        ctx = C.c_void_p()
        N.check(N.lib().supra_bf_create(C.byref(self._gc), C.byref(ctx)))
        self._ctx = ctx

    def close(self) -> None:
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            N.lib().supra_bf_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def perform_rx_beamforming(self, sample_beamformer: int, raw: USRawData, f_number: float,
                               speed_of_sound_mm_per_s: float, window_type: int, window_parameter: float,
                               interpolate_transmits: bool, additional_offset: int,
                               out_dtype=None, out=None, stream=None) -> USImage:
        """RxBeamformerCuda::performRxBeamforming (RxBeamformerCuda.cu:539-663)."""
        data = raw.data
        frames = 1 if data.dim() == 3 else int(data.shape[0])
        if data.dim() not in (3, 4) or not data.is_cuda or not data.is_contiguous():
            raise ValueError("raw.data must be a contiguous device tensor [tx][ch][s] or [frame][tx][ch][s]")
        out_dtype = out_dtype or torch.float32
        g = self.geometry
        shape = (g.num_depths, g.num_rx) if data.dim() == 3 else (frames, g.num_depths, g.num_rx)
        if out is None:
            out = torch.empty(shape, dtype=out_dtype, device=data.device)
        elif tuple(out.shape) != shape or out.dtype != out_dtype or not out.is_contiguous():
            raise ValueError(f"out must be a contiguous {out_dtype} tensor of shape {shape}")
        N.check(N.lib().supra_das_run_frames(
            self._ctx, int(sample_beamformer), frames, C.c_void_p(data.data_ptr()), _dtype_code(data.dtype),
            raw.num_scanlines, raw.num_received_channels, raw.num_samples, float(raw.sampling_frequency),
            float(f_number), float(speed_of_sound_mm_per_s), int(window_type), float(window_parameter),
            int(bool(interpolate_transmits)), int(additional_offset), C.c_void_p(out.data_ptr()),
            _dtype_code(out_dtype), C.c_void_p(_stream_handle(stream))))
        return USImage(out, (g.num_rx, g.num_depths), {"imageState": "RF", "geometry": g})


class _Param:
    def __init__(self, default, lo=None, hi=None, choices=None):
        self.default, self.lo, self.hi, self.choices = default, lo, hi, choices

    def validate(self, key, value):
        if self.choices is not None:
            if value not in self.choices:
                raise ValueError(f"{key}: {value!r} not in {sorted(self.choices)}")
        elif self.lo is not None and not (self.lo <= value <= self.hi):
            raise ValueError(f"{key}: {value} outside [{self.lo}, {self.hi}]")
        return value


class BeamformingNode:
    """BeamformingNode (BeamformingNode.cpp:24-232): XML parameters and dispatch."""

    PARAMETERS = {
        "fNumber": _Param(1.0, 0.1, 4.0),
        "windowType": _Param("Rectangular", choices=set(WINDOW_TYPES)),
        "windowParameter": _Param(0.0, 0.0, 10.0),
        "speedOfSound": _Param(1540.0, 1000.0, 2000.0),
        "beamformerType": _Param("DelayAndSum", choices=set(BEAMFORMER_TYPES)),
        "interpolateTransmits": _Param(False, choices={False, True}),
        "additionalOffset": _Param(0, -1000, 1000),
        "outputType": _Param("float", choices=set(OUTPUT_TYPES)),
    }
    # the spelling rxBeamformerType is accepted as an alias of the beamformer key
    ALIASES = {"rxBeamformerType": "beamformerType"}

    def __init__(self, node_id: str = "BeamformingNode", **params):
        self.node_id = node_id
        self.config = {k: p.default for k, p in self.PARAMETERS.items()}
        self._engine: RxBeamformer | None = None
        self._engine_geometry = None
        for k, v in params.items():
            self.set(k, v)

    def set(self, key: str, value) -> None:
        key = self.ALIASES.get(key, key)
        if key not in self.PARAMETERS:
            raise KeyError(f"BeamformingNode has no parameter {key!r}")
        self.config[key] = self.PARAMETERS[key].validate(key, value)

    @classmethod
    def from_xml(cls, xml_text: str) -> "BeamformingNode":
        """Parse a <node type="BeamformingNode"> element with <param name type> children
        (ConfigurationDictionary.cpp:24-113)."""
        import xml.etree.ElementTree as ET
        root = ET.fromstring(xml_text)
        node = root if root.tag == "node" else root.find(".//node")
        obj = cls(node.get("id", "BeamformingNode"))
        for p in node.findall("param"):
            name, typ, text = p.get("name"), p.get("type"), (p.text or "").strip()
            if typ in ("double", "float"):
                val = float(text)
            elif typ in ("int", "int32_t", "uint32_t"):
                val = int(text)
            elif typ == "bool":
                val = text.lower() in ("1", "true")
            else:
                val = text
            obj.set(name, val)
        return obj

    def process(self, raw: USRawData, stream=None) -> USImage:
        """checkTypeAndBeamform (BeamformingNode.cpp:128-177)."""
        if self._engine is None or self._engine_geometry is not raw.geometry:
            self._engine = RxBeamformer(raw.geometry)
            self._engine_geometry = raw.geometry
        bt = BEAMFORMER_TYPES[self.config["beamformerType"]]
        out_dtype = torch.float32 if self.config["outputType"] == "float" else torch.int16
        img = self._engine.perform_rx_beamforming(
            bt, raw, self.config["fNumber"], self.config["speedOfSound"] * 1000.0,
            WINDOW_TYPES[self.config["windowType"]], self.config["windowParameter"],
            self.config["interpolateTransmits"], self.config["additionalOffset"], out_dtype=out_dtype,
            stream=stream)
        img.properties.update({
            "Beamformer.FNumber": self.config["fNumber"],
            "Beamformer.WindowType": self.config["windowType"],
            "Beamformer.WindowParameter": self.config["windowParameter"],
            "Beamformer.RxSpeedOfSound": self.config["speedOfSound"] * 1000.0,
            "Beamformer.BeamformerType": self.config["beamformerType"],
            "Beamformer.InterpolateTransmits": self.config["interpolateTransmits"],
        })
        return img


class RawDelay:
    """RawDelay (src/SupraLib/Beamformer/RawDelay.h:38-48): per-channel delayed,
    apodised RF [scanline][channel][depth] for the MV beamformer."""

    def __init__(self, geometry: RxGeometry):
        self._engine = RxBeamformer(geometry)
        self.geometry = geometry

    def close(self) -> None:
        self._engine.close()

    def perform_delay(self, raw: USRawData, f_number: float, speed_of_sound_mm_per_s: float, window_type: int,
                      window_parameter: float, out_dtype=None, out=None, stream=None) -> USRawData:
        """RawDelay::performDelay (RawDelay.cu:555-650)."""
        data = raw.data
        if data.dim() != 3 or not data.is_cuda or not data.is_contiguous():
            raise ValueError("raw.data must be a contiguous device tensor [tx][ch][s]")
        out_dtype = out_dtype or torch.float32
        g = self.geometry
        shape = (g.num_rx, raw.num_received_channels, g.num_depths)
        if out is None:
            out = torch.empty(shape, dtype=out_dtype, device=data.device)
        elif tuple(out.shape) != shape or out.dtype != out_dtype or not out.is_contiguous():
            raise ValueError(f"out must be a contiguous {out_dtype} tensor of shape {shape}")
        fs_out = C.c_double()
        N.check(N.lib().supra_rawdelay_run(
            self._engine._ctx, C.c_void_p(data.data_ptr()), _dtype_code(data.dtype), raw.num_scanlines,
            raw.num_received_channels, raw.num_samples, float(raw.sampling_frequency), float(f_number),
            float(speed_of_sound_mm_per_s), int(window_type), float(window_parameter), C.c_void_p(out.data_ptr()),
            _dtype_code(out_dtype), C.byref(fs_out), C.c_void_p(_stream_handle(stream))))
        return USRawData(out, fs_out.value, g, raw.num_elements, raw.element_layout)


def mv_perform_rx_beamforming(raw: USRawData, sub_array_size: int = 0, temporal_smoothing: int = 3,
                              sub_array_scaling_power: float = 1.5, compute_means: bool = False,
                              loading_factor: float = 1.0, out_dtype=None, out=None, stream=None) -> USImage:
    """RxBeamformerMV::performRxBeamforming (RxBeamformerMV.cu:382-614) on
    pre-delayed data [scanline][channel][sample]; returns [sample][scanline]."""
    data = raw.data
    if data.dim() != 3 or not data.is_cuda or not data.is_contiguous():
        raise ValueError("raw.data must be a contiguous device tensor [scanline][channel][sample]")
    n_rx, n_ch, n_s = (int(v) for v in data.shape)
    out_dtype = out_dtype or torch.float32
    if out is None:
        out = torch.empty((n_s, n_rx), dtype=out_dtype, device=data.device)
    elif tuple(out.shape) != (n_s, n_rx) or out.dtype != out_dtype or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous {out_dtype} tensor of shape {(n_s, n_rx)}")
    N.check(N.lib().supra_mv_run(C.c_void_p(data.data_ptr()), _dtype_code(data.dtype), n_rx, n_ch, n_s,
                                 int(sub_array_size), int(temporal_smoothing), float(sub_array_scaling_power),
                                 int(bool(compute_means)), float(loading_factor), C.c_void_p(out.data_ptr()),
                                 _dtype_code(out_dtype), C.c_void_p(_stream_handle(stream))))
    return USImage(out, (n_rx, n_s), {"imageState": "RF"})


def mvpcg_perform_rx_beamforming(raw: USRawData, sub_array_size: int = 0, temporal_smoothing: int = 3,
                                 max_iterations_override: int = 0, convergence_threshold: float = 1e-4,
                                 sub_array_scaling_power: float = 1.5, compute_means: bool = False,
                                 out_dtype=None, out=None, stream=None) -> USImage:
    """RxBeamformerMVpcg::performRxBeamforming (RxBeamformerMVpcg.cu:628-800):
    minimum variance with the weights from Jacobi-preconditioned conjugate
    gradients, on pre-delayed data [scanline][channel][sample]."""
    data = raw.data
    if data.dim() != 3 or not data.is_cuda or not data.is_contiguous():
        raise ValueError("raw.data must be a contiguous device tensor [scanline][channel][sample]")
    n_rx, n_ch, n_s = (int(v) for v in data.shape)
    out_dtype = out_dtype or torch.float32
    if out is None:
        out = torch.empty((n_s, n_rx), dtype=out_dtype, device=data.device)
    elif tuple(out.shape) != (n_s, n_rx) or out.dtype != out_dtype or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous {out_dtype} tensor of shape {(n_s, n_rx)}")
    N.check(N.lib().supra_mvpcg_run(C.c_void_p(data.data_ptr()), _dtype_code(data.dtype), n_rx, n_ch, n_s,
                                    int(sub_array_size), int(temporal_smoothing), int(max_iterations_override),
                                    float(convergence_threshold), float(sub_array_scaling_power),
                                    int(bool(compute_means)), C.c_void_p(out.data_ptr()), _dtype_code(out_dtype),
                                    C.c_void_p(_stream_handle(stream))))
    return USImage(out, (n_rx, n_s), {"imageState": "RF"})


def mv_das_perform_rx_beamforming(engine: "RxBeamformer", raw: USRawData, f_number: float,
                                  speed_of_sound_mm_per_s: float, window_type: int, window_parameter: float,
                                  sub_array_size: int = 0, temporal_smoothing: int = 3,
                                  sub_array_scaling_power: float = 1.5, compute_means: bool = False,
                                  loading_factor: float = 1.0, out_dtype=None, out=None, stream=None) -> USImage:
    """RawDelay::performDelay followed by RxBeamformerMV::performRxBeamforming
    in one fused launch (supra_mv_das_run): RF [tx][ch][s] -> [depth][scanline]."""
    data = raw.data
    if data.dim() != 3 or not data.is_cuda or not data.is_contiguous():
        raise ValueError("raw.data must be a contiguous device tensor [tx][ch][s]")
    g = engine.geometry
    out_dtype = out_dtype or torch.float32
    if out is None:
        out = torch.empty((g.num_depths, g.num_rx), dtype=out_dtype, device=data.device)
    elif tuple(out.shape) != (g.num_depths, g.num_rx) or out.dtype != out_dtype or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous {out_dtype} tensor of shape {(g.num_depths, g.num_rx)}")
    N.check(N.lib().supra_mv_das_run(
        engine._ctx, C.c_void_p(data.data_ptr()), _dtype_code(data.dtype), raw.num_scanlines,
        raw.num_received_channels, raw.num_samples, float(raw.sampling_frequency), float(f_number),
        float(speed_of_sound_mm_per_s), int(window_type), float(window_parameter), int(sub_array_size),
        int(temporal_smoothing), float(sub_array_scaling_power), int(bool(compute_means)), float(loading_factor),
        C.c_void_p(out.data_ptr()), _dtype_code(out_dtype), C.c_void_p(_stream_handle(stream))))
    return USImage(out, (g.num_rx, g.num_depths), {"imageState": "RF", "geometry": g})


class RawDelayNode(BeamformingNode):
    """RawDelayNode (RawDelayNode.cpp:25-195): parameters RawDelayNode.cpp:43-47."""

    PARAMETERS = {
        "fNumber": _Param(1.0, 0.01, 4.0),
        "windowType": _Param("Rectangular", choices=set(WINDOW_TYPES)),
        "windowParameter": _Param(0.0, 0.0, 10.0),
        "speedOfSound": _Param(1540.0, 1000.0, 2000.0),
        "outputType": _Param("float", choices=set(OUTPUT_TYPES)),
    }
    ALIASES = {}

    def process(self, raw: USRawData, stream=None) -> USRawData:
        """checkTypeAndDelay (RawDelayNode.cpp:108-163)."""
        if self._engine is None or self._engine_geometry is not raw.geometry:
            self._engine = RawDelay(raw.geometry)
            self._engine_geometry = raw.geometry
        out_dtype = torch.float32 if self.config["outputType"] == "float" else torch.int16
        return self._engine.perform_delay(raw, self.config["fNumber"], self.config["speedOfSound"] * 1000.0,
                                          WINDOW_TYPES[self.config["windowType"]], self.config["windowParameter"],
                                          out_dtype=out_dtype, stream=stream)


class BeamformingMVNode(BeamformingNode):
    """BeamformingMVNode (BeamformingMVNode.cpp:25-172): parameters :42-46 plus
    `loadingFactor` (diagonal loading = loadingFactor * trace / L; 1 = reference)."""

    PARAMETERS = {
        "subArraySize": _Param(0, 0, 64),
        "temporalSmoothing": _Param(3, 0, 10),
        "outputType": _Param("float", choices=set(OUTPUT_TYPES)),
        "subArrayScalingPower": _Param(1.5, 0.5, 3.0),
        "computeMeans": _Param(False, choices={False, True}),
        "loadingFactor": _Param(1.0, 0.0, 1e12),
    }
    ALIASES = {}

    def process(self, raw: USRawData, stream=None) -> USImage:
        """checkTypeAndBeamform (BeamformingMVNode.cpp:121-163); input must be pre-delayed."""
        out_dtype = torch.float32 if self.config["outputType"] == "float" else torch.int16
        return mv_perform_rx_beamforming(raw, self.config["subArraySize"], self.config["temporalSmoothing"],
                                         self.config["subArrayScalingPower"], self.config["computeMeans"],
                                         self.config["loadingFactor"], out_dtype=out_dtype, stream=stream)


class BeamformingMVpcgNode(BeamformingNode):
    """BeamformingMVpcgNode (BeamformingMVpcgNode.cpp:38-193): parameters
    :55-61 (the convergence threshold is 10^convergenceThresholdExponent)."""

    PARAMETERS = {
        "subArraySize": _Param(0, 0, 64),
        "temporalSmoothing": _Param(3, 0, 10),
        "outputType": _Param("float", choices=set(OUTPUT_TYPES)),
        "maxIterationsOverride": _Param(0, 0, 10000),
        "convergenceThresholdExponent": _Param(-4.0, -100.0, 0.0),
        "subArrayScalingPower": _Param(1.5, 0.5, 3.0),
        "computeMeans": _Param(False, choices={False, True}),
    }
    ALIASES = {}

    def process(self, raw: USRawData, stream=None) -> USImage:
        """checkTypeAndBeamform (BeamformingMVpcgNode.cpp:147-190); input must be pre-delayed."""
        out_dtype = torch.float32 if self.config["outputType"] == "float" else torch.int16
        return mvpcg_perform_rx_beamforming(
            raw, self.config["subArraySize"], self.config["temporalSmoothing"],
            self.config["maxIterationsOverride"], 10.0 ** self.config["convergenceThresholdExponent"],
            self.config["subArrayScalingPower"], self.config["computeMeans"], out_dtype=out_dtype, stream=stream)


def window_lut(window_type: int, window_parameter: float) -> np.ndarray:
    """The library's window table (supra_window_lut)."""
    out = np.zeros(N.WINDOW_NUM_ENTRIES, dtype=np.float32)
    N.check(N.lib().supra_window_lut(int(window_type), float(window_parameter), N.ptr(out, C.c_float)))
    return out
