"""Receive geometry (RxBeamformerParameters) on the host, and the
transmit/scan configuration that produces it (Beamformer + USTransducer).

``BeamformerConfig`` mirrors the setters of the reference's Beamformer
(src/SupraLib/Beamformer/Beamformer.cpp:91-360); ``RxGeometry.from_config``
runs the C++ geometry producer (supra_geometry_build) and keeps the tables as
numpy arrays, in millimetres and seconds like RxBeamformerParameters.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

# numpy view of supra_scanline_rx (same layout as the ctypes structure)
SCANLINE_DTYPE = np.dtype({
    "names": ["position", "direction", "tx_weights", "tx", "max_element_distance"],
    "formats": [("<f8", 3), ("<f8", 3), ("<f8", 4),
                (np.dtype({"names": ["first", "last", "tx_idx", "reserved", "initial_delay"],
                           "formats": [("<u2", 2), ("<u2", 2), "<u2", ("<u2", 3), "<f8"],
                           "offsets": [0, 4, 8, 10, 16], "itemsize": 24}), 4),
                ("<f8", 2)],
    "offsets": [0, 24, 48, 80, 176],
    "itemsize": 192,
})
assert SCANLINE_DTYPE.itemsize == C.sizeof(N.ScanlineRx)

_SCAN = {"linear": N.SCAN_LINEAR, "phased": N.SCAN_PHASED, "biphased": N.SCAN_BIPHASED,
         "planewave": N.SCAN_PLANEWAVE}
_TRANSDUCER = {"linear": N.TRANSDUCER_LINEAR, "planar": N.TRANSDUCER_PLANAR}


@dataclass
class BeamformerConfig:
    """Scan and transducer parameters (Beamformer.h / USTransducer.h)."""
    scan_type: str = "linear"            # linear | phased | biphased | planewave
    transducer_type: str = "linear"      # linear | planar
    element_layout: tuple = (128, 1)
    pitch: tuple = (0.3, 0.3)            # mm
    num_scanlines: tuple = (128, 1)      # transmit events
    rx_scanline_subdivision: tuple = (1, 1)
    max_aperture: tuple = (64, 1)
    tx_max_aperture: tuple = (64, 1)
    depth: float = 40.0                  # mm
    num_depths: int = 2080
    steering_angle_deg: tuple = (0.0, 0.0)
    sector_angle_deg: tuple = (0.0, 0.0)
    tx_focus_active: bool = True
    tx_focus_depth: float = 20.0         # mm
    tx_focus_width: float = 0.0          # mm
    speed_of_sound: float = 1540.0       # m/s

    def to_c(self) -> N.BeamformerConfigC:
        c = N.BeamformerConfigC()
        c.scan_type = _SCAN[self.scan_type]
        c.transducer_type = _TRANSDUCER[self.transducer_type]
        c.element_layout[:] = list(self.element_layout)
        c.pitch[:] = list(self.pitch)
        c.num_scanlines[:] = list(self.num_scanlines)
        c.rx_scanline_subdivision[:] = list(self.rx_scanline_subdivision)
        c.max_aperture[:] = list(self.max_aperture)
        c.tx_max_aperture[:] = list(self.tx_max_aperture)
        c.depth = self.depth
        c.num_depths = self.num_depths
        c.steering_angle_deg[:] = list(self.steering_angle_deg)
        c.sector_angle_deg[:] = list(self.sector_angle_deg)
        c.tx_focus_active = int(self.tx_focus_active)
        c.tx_focus_depth = self.tx_focus_depth
        c.tx_focus_width = self.tx_focus_width
        c.speed_of_sound = self.speed_of_sound
        return c


@dataclass
class RxGeometry:
    """RxBeamformerParameters (RxBeamformerParameters.h:38-112) as numpy arrays."""
    scanlines: np.ndarray                 # SCANLINE_DTYPE [num_rx]
    rx_layout: tuple
    depths: np.ndarray                    # float32 [num_depths], mm
    element_x: np.ndarray                 # float32 [num_elements], mm
    element_y: np.ndarray                 # float32 [num_elements], mm
    element_layout: tuple
    speed_of_sound_mm_per_s: float
    depth: float
    element_to_channel: np.ndarray | None = None
    _keep: list = field(default_factory=list, repr=False)

    @property
    def num_rx(self) -> int:
        return int(self.scanlines.shape[0])

    @property
    def num_depths(self) -> int:
        return int(self.depths.shape[0])

    @property
    def num_elements(self) -> int:
        return int(self.element_x.shape[0])

    @property
    def is3d(self) -> bool:
        return self.rx_layout[0] > 1 and self.rx_layout[1] > 1

    @classmethod
    def from_config(cls, cfg: BeamformerConfig) -> "RxGeometry":
        lib = N.lib()
        cc = cfg.to_c()
        n_rx, n_el = C.c_uint32(), C.c_uint32()
        lay = (C.c_uint32 * 2)()
        N.check(lib.supra_geometry_sizes(C.byref(cc), C.byref(n_rx), lay, C.byref(n_el)))
        scan = np.zeros(n_rx.value, dtype=SCANLINE_DTYPE)
        depths = np.zeros(cfg.num_depths, dtype=np.float32)
        ex = np.zeros(n_el.value, dtype=np.float32)
        ey = np.zeros(n_el.value, dtype=np.float32)
        N.check(lib.supra_geometry_build(C.byref(cc), scan.ctypes.data_as(C.POINTER(N.ScanlineRx)),
                                         N.ptr(depths, C.c_float), N.ptr(ex, C.c_float), N.ptr(ey, C.c_float)))
        el_layout = (cfg.element_layout[0], 1 if cfg.transducer_type == "linear" else cfg.element_layout[1])
        return cls(scan, (lay[0], lay[1]), depths, ex, ey, el_layout, cfg.speed_of_sound * 1000.0, cfg.depth)

    def to_c(self) -> N.RxGeometryC:
        """C struct pointing into this object's arrays (valid while self lives)."""
        g = N.RxGeometryC()
        g.num_rx_scanlines = self.num_rx
        g.rx_scanline_layout[:] = list(self.rx_layout)
        g.num_depths = self.num_depths
        g.num_elements = self.num_elements
        g.element_layout[:] = list(self.element_layout)
        self.scanlines = np.ascontiguousarray(self.scanlines)
        self.depths = np.ascontiguousarray(self.depths, dtype=np.float32)
        self.element_x = np.ascontiguousarray(self.element_x, dtype=np.float32)
        self.element_y = np.ascontiguousarray(self.element_y, dtype=np.float32)
        g.scanlines = self.scanlines.ctypes.data_as(C.POINTER(N.ScanlineRx))
        g.depths = N.ptr(self.depths, C.c_float)
        g.element_x = N.ptr(self.element_x, C.c_float)
        g.element_y = N.ptr(self.element_y, C.c_float)
        if self.element_to_channel is not None:
            self.element_to_channel = np.ascontiguousarray(self.element_to_channel, dtype=np.int32)
            g.element_to_channel = N.ptr(self.element_to_channel, C.c_int32)
        g.speed_of_sound_mm_per_s = self.speed_of_sound_mm_per_s
        g.depth = self.depth
        return g

    def transmit_tables(self, num_tx: int):
        """Per-transmit origin/direction/t0/aperture for the RF synthesizer,
        taken from the receive scanlines whose first transmit carries weight."""
        origin = np.zeros((num_tx, 3))
        direction = np.zeros((num_tx, 3))
        direction[:, 1] = 1.0
        t0 = np.zeros(num_tx)
        first = np.zeros(num_tx, dtype=np.uint32)
        last = np.zeros(num_tx, dtype=np.uint32)
        seen = np.zeros(num_tx, dtype=bool)
        lx = self.element_layout[0]
        for s in self.scanlines:
            for k in range(4):
                if s["tx_weights"][k] <= 0:
                    continue
                tp = s["tx"][k]
                i = int(tp["tx_idx"])
                if i >= num_tx or seen[i]:
                    continue
                seen[i] = True
                origin[i] = s["position"]
                direction[i] = s["direction"]
                t0[i] = tp["initial_delay"]
                first[i] = int(tp["first"][0]) + int(tp["first"][1]) * lx
                last[i] = int(tp["last"][0]) + int(tp["last"][1]) * lx
        return origin, direction, t0, first, last

    def element_positions(self) -> np.ndarray:
        """[num_elements, 3] (x, depth=0, elevation) in mm."""
        pos = np.zeros((self.num_elements, 3))
        pos[:, 0] = self.element_x
        pos[:, 2] = self.element_y
        return pos
