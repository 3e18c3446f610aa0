"""Scan conversion of beamformed images onto a Cartesian grid, mirroring
SUPRA's ScanConverter and ScanConverterNode:

* ``ScanConverter`` — ScanConverter (ScanConverter.h:25-93):
  ``update_internals(geometry, num_samples, resolution)`` builds the pixel
  tables (updateInternals, ScanConverter.cu:516-810), ``convert(image)``
  resamples one image or a batch of frames (convert, :436-488),
  ``get_mask()`` the uint8 mask (:430-433), ``image_size`` (getImageSize);
* ``ScanConverterNode`` — the node's XML parameters with the reference's
  names, ranges and defaults (ScanConverterNode.cpp:41-43) and its rule of
  rebuilding the tables only when the geometry, the sample count or the
  resolution changes (:118-182).

All computation is in libsupra_bf.so (supra_scanconv_*); there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native as N
from .beamformer import USImage, _Param, _stream_handle
from .geometry import RxGeometry

_CODES = {torch.int16: N.SUPRA_INT16, torch.float32: N.SUPRA_FLOAT32, torch.uint8: N.SUPRA_UINT8}
OUTPUT_TYPES = {"float": torch.float32, "int16": torch.int16, "uint8": torch.uint8}
NORMALIZED_DEPTH_WEIGHT = 1


def _code(dtype) -> int:
    if dtype not in _CODES:
        raise TypeError(f"unsupported element type {dtype}")
    return _CODES[dtype]


class ScanConverter:
    """Tables for one scan geometry, resident on the current device."""

    def __init__(self):
        self._sc = None
        self._key = None

    def close(self) -> None:
        if self._sc is not None and self._sc.value:
            N.lib().supra_scanconv_destroy(self._sc)
        self._sc = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update_internals(self, geometry: RxGeometry, num_samples: int | None = None, resolution: float = 0.0,
                         normalized_depth_weight: bool = False) -> None:
        """Build the tables for `geometry` (scanline positions/directions,
        layout, imaging depth); num_samples defaults to the geometry's depth
        count; resolution <= 0 means the sample distance."""
        num_samples = int(num_samples if num_samples is not None else geometry.num_depths)
        scan = np.ascontiguousarray(geometry.scanlines)
        layout = (C.c_uint32 * 2)(*(int(v) for v in geometry.rx_layout))
        sc = C.c_void_p()
        N.check(N.lib().supra_scanconv_create(scan.ctypes.data, layout, float(geometry.depth), num_samples,
                                              float(resolution),
                                              NORMALIZED_DEPTH_WEIGHT if normalized_depth_weight else 0,
                                              C.byref(sc)))
        self.close()
        self._sc = sc
        self._layout = tuple(int(v) for v in geometry.rx_layout)
        self._num_samples = num_samples
        size = (C.c_uint32 * 3)()
        lo, hi = (C.c_double * 3)(), (C.c_double * 3)()
        res = C.c_double()
        N.check(N.lib().supra_scanconv_geometry(sc, size, lo, hi, C.byref(res)))
        self.image_size = tuple(int(v) for v in size)
        self.bb_min, self.bb_max = tuple(lo), tuple(hi)
        self.resolution = float(res.value)

    @property
    def is_3d(self) -> bool:
        return not (self._layout[0] == 1 or self._layout[1] == 1)

    @property
    def num_pixels(self) -> int:
        return int(np.prod(self.image_size))

    def _require(self):
        if self._sc is None:
            raise RuntimeError("ScanConverter: update_internals() has not been called")

    def tables(self):
        """Host copies: (mask uint8, sample index uint32, weights float32 [3][pixels])."""
        self._require()
        n = self.num_pixels
        mask = np.empty(n, np.uint8)
        idx = np.empty(n, np.uint32)
        w = np.empty((3, n), np.float32)
        N.check(N.lib().supra_scanconv_tables(self._sc, mask.ctypes.data, idx.ctypes.data, w.ctypes.data))
        return mask, idx, w

    def _shape(self, frames=None) -> tuple:
        sx, sy, sz = self.image_size
        img = (sz, sy, sx) if self.is_3d else (sy, sx)
        return img if frames is None else (frames, *img)

    def get_mask(self, stream=None) -> torch.Tensor:
        self._require()
        m = torch.empty(self._shape(), dtype=torch.uint8, device="cuda")
        N.check(N.lib().supra_scanconv_mask(self._sc, C.c_void_p(m.data_ptr()), C.c_void_p(_stream_handle(stream))))
        return m

    def convert(self, image, out_dtype=torch.float32, out=None, stream=None) -> torch.Tensor:
        """[num_samples][scanlines] (or [frames][num_samples][scanlines]) ->
        [y][x] (2D) / [z][y][x] (3D), with a leading frame axis for batches."""
        self._require()
        data = image.data if isinstance(image, USImage) else image
        n_rx = self._layout[0] * self._layout[1]
        if not data.is_cuda or not data.is_contiguous():
            raise ValueError("image must be a contiguous device tensor")
        if data.dim() == 2:
            frames, batched = 1, False
        elif data.dim() == 3:
            frames, batched = int(data.shape[0]), True
        else:
            raise ValueError("image must be [samples][scanlines] or [frames][samples][scanlines]")
        if tuple(data.shape[-2:]) != (self._num_samples, n_rx):
            raise ValueError(f"image shape {tuple(data.shape)} does not match the tables "
                             f"({self._num_samples} samples x {n_rx} scanlines)")
        shape = self._shape(frames if batched else None)
        if out is None:
            out = torch.empty(shape, dtype=out_dtype, device=data.device)
        elif tuple(out.shape) != shape or out.dtype != out_dtype or not out.is_contiguous():
            raise ValueError(f"out must be a contiguous {out_dtype} tensor of shape {shape}")
        N.check(N.lib().supra_scanconv_run(self._sc, C.c_void_p(data.data_ptr()), _code(data.dtype), frames,
                                           C.c_void_p(out.data_ptr()), _code(out_dtype),
                                           C.c_void_p(_stream_handle(stream))))
        return out


class ScanConverterNode:
    """ScanConverterNode (ScanConverterNode.cpp:24-200)."""

    PARAMETERS = {
        "imageResolution": _Param(0.1, 0.01, 5.0),
        "imageResolutionForced": _Param(False, choices={False, True}),
        "outputType": _Param("float", choices=set(OUTPUT_TYPES)),
    }

    def __init__(self, node_id: str = "ScanConverterNode", **params):
        self.node_id = node_id
        self.config = {k: p.default for k, p in self.PARAMETERS.items()}
        for k, v in params.items():
            self.set(k, v)
        self.converter = ScanConverter()
        self._seen = None
        self._needs_update = True
        self.mask = None

    def set(self, key: str, value) -> None:
        if key not in self.PARAMETERS:
            raise KeyError(f"{self.node_id}: unknown parameter {key}")
        self.config[key] = self.PARAMETERS[key].validate(key, value)
        if key != "outputType":   # configurationEntryChanged (:55-71)
            self._needs_update = True

    def process(self, image: USImage, geometry: RxGeometry | None = None, stream=None) -> USImage:
        geometry = geometry if geometry is not None else image.properties.get("geometry")
        if geometry is None:
            raise ValueError(f"{self.node_id}: the image carries no scanline geometry")
        data = image.data
        num_samples = int(data.shape[-2])
        key = (id(geometry), num_samples)
        if self._needs_update or key != self._seen:
            res = float(self.config["imageResolution"]) if self.config["imageResolutionForced"] else 0.0
            self.converter.update_internals(geometry, num_samples, res)
            self._seen = key
            self._needs_update = False
            self.mask = self.converter.get_mask(stream)   # sendMask (:73-86), once per update
        out = self.converter.convert(data, OUTPUT_TYPES[self.config["outputType"]], stream=stream)
        props = dict(image.properties)
        props.update({"imageState": "Scan", "imageResolution": self.converter.resolution})
        return USImage(out, self.converter.image_size, props)
