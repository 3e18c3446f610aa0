"""SUPRA's raw-data mock files (recorded acquisitions) through the C ABI:

* ``read_mock(path)`` — RxBeamformerParameters::readMetaDataForMock
  (RxBeamformerParameters.cpp:158-185): the ``RxGeometry`` and the sizes of a
  ``.json`` or ASCII ``.mock`` metadata file;
* ``write_mock(path, geometry, meta)`` — writeMetaDataForMock (:93-156) for
  ``.json``, or the ASCII v1 layout for ``.mock``;
* ``MockRawDataSource`` — UltrasoundInterfaceRawDataMock
  (InputOutput/UltrasoundInterfaceRawDataMock.cpp:49-204): int16 frames
  [tx][channel][sample] from one or more raw files, uploaded to the GPU as
  ``USRawData``; sequences loop unless ``stream_sequence_once``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .beamformer import USRawData
from .geometry import SCANLINE_DTYPE, RxGeometry


@dataclass
class MockMeta:
    num_elements: int
    element_layout: tuple
    num_channels: int
    num_samples: int
    num_tx: int
    rx_scanline_layout: tuple
    num_depths: int
    depth: float
    sampling_frequency_hz: float
    speed_of_sound_mm_per_s: float

    def to_c(self) -> N.MockMetaC:
        m = N.MockMetaC()
        m.num_elements = self.num_elements
        m.element_layout[:] = list(self.element_layout)
        m.num_channels, m.num_samples, m.num_tx = self.num_channels, self.num_samples, self.num_tx
        m.rx_scanline_layout[:] = list(self.rx_scanline_layout)
        m.num_depths = self.num_depths
        m.depth, m.sampling_frequency_hz = self.depth, self.sampling_frequency_hz
        m.speed_of_sound_mm_per_s = self.speed_of_sound_mm_per_s
        return m

    @classmethod
    def from_c(cls, m: N.MockMetaC) -> "MockMeta":
        return cls(m.num_elements, tuple(m.element_layout), m.num_channels, m.num_samples, m.num_tx,
                   tuple(m.rx_scanline_layout), m.num_depths, m.depth, m.sampling_frequency_hz,
                   m.speed_of_sound_mm_per_s)

    @property
    def frame_shape(self) -> tuple:
        return (self.num_tx, self.num_channels, self.num_samples)


def read_mock(path: str):
    """(RxGeometry, MockMeta) of a .json or .mock metadata file."""
    lib = N.lib()
    m = N.MockMetaC()
    N.check(lib.supra_mock_read_meta(str(path).encode(), C.byref(m)))
    meta = MockMeta.from_c(m)
    n_rx = meta.rx_scanline_layout[0] * meta.rx_scanline_layout[1]
    scan = np.zeros(n_rx, dtype=SCANLINE_DTYPE)
    depths = np.zeros(meta.num_depths, dtype=np.float32)
    ex = np.zeros(meta.num_elements, dtype=np.float32)
    ey = np.zeros(meta.num_elements, dtype=np.float32)
    N.check(lib.supra_mock_read_geometry(str(path).encode(), scan.ctypes.data, depths.ctypes.data, ex.ctypes.data,
                                         ey.ctypes.data))
    geom = RxGeometry(scan, tuple(meta.rx_scanline_layout), depths, ex, ey, tuple(meta.element_layout),
                      meta.speed_of_sound_mm_per_s, meta.depth)
    return geom, meta


def write_mock(path: str, geometry: RxGeometry, meta: MockMeta) -> None:
    """Write the metadata as .json (writeMetaDataForMock) or ASCII .mock v1."""
    gc = geometry.to_c()
    m = meta.to_c()
    fn = N.lib().supra_mock_write_json if str(path).endswith(".json") else N.lib().supra_mock_write_ascii
    N.check(fn(str(path).encode(), C.byref(m), C.byref(gc)))


def write_frames(path: str, frames: np.ndarray) -> None:
    """Append-free write of int16 frames [frame][tx][ch][s] as one raw file."""
    np.ascontiguousarray(frames, dtype=np.int16).tofile(str(path))


class MockRawDataSource:
    """Frames of one or more raw files as USRawData on the GPU, in order,
    looping over the sequences (UltrasoundInterfaceRawDataMock::readNextFrame)."""

    def __init__(self, meta_path: str, raw_paths, stream_sequence_once: bool = False, device: str = "cuda"):
        import torch
        self._torch = torch
        self.geometry, self.meta = read_mock(meta_path)
        self.raw_paths = [p.strip() for p in (raw_paths.split(",") if isinstance(raw_paths, str) else raw_paths)]
        self.stream_sequence_once = stream_sequence_once
        self.device = device
        self._m = self.meta.to_c()
        self.lengths = []
        for p in self.raw_paths:
            n = C.c_uint64()
            N.check(N.lib().supra_mock_count_frames(p.encode(), C.byref(self._m), C.byref(n)))
            self.lengths.append(int(n.value))
        self._seq, self._frame, self.done = 0, 0, False

    def read_frame(self, sequence: int, frame: int) -> np.ndarray:
        out = np.empty(self.meta.frame_shape, dtype=np.int16)
        N.check(N.lib().supra_mock_read_frames(self.raw_paths[sequence].encode(), C.byref(self._m), frame, 1,
                                               out.ctypes.data))
        return out

    def next(self) -> USRawData | None:
        if self.done:
            return None
        host = self.read_frame(self._seq, self._frame)
        self._frame = (self._frame + 1) % self.lengths[self._seq]
        if self._frame == 0:
            self._seq = (self._seq + 1) % len(self.lengths)
            if self._seq == 0 and self.stream_sequence_once:
                self.done = True
        data = self._torch.from_numpy(host).to(self.device)
        return USRawData(data, self.meta.sampling_frequency_hz, self.geometry, self.meta.num_elements,
                         tuple(self.meta.element_layout))
