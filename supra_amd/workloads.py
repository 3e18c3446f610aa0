"""Seeded acquisition workloads: scan geometry, beamforming parameters and
synthetic RF.

The named workloads are the benchmark configurations of this beamformer:

* ``linear_2d``  — 128-element linear array (pitch 0.3 mm), c = 1540 m/s,
  fs = 40 MHz, 128 focused transmits (focus 20 mm), 64-channel receive
  aperture (channel = element mod 64), F# 1, Hann, 2080 samples and 2080
  depths over 40 mm; 200 scatterers in 38 x 40 mm, 7 MHz pulse with 60 %
  bandwidth, 30 dB SNR, int16 RF scaled to peak 16384.
* ``matrix_3d`` — 32 x 32 matrix array (pitch 0.3 mm), fs = 20 MHz,
  3.5 MHz pulse, 64 x 64 focused biphased scanlines fanning over a 40 mm
  cube (sector 2*atan(20/40) per axis), 1024 channels, 1000 samples and
  1000 depths, full 32 x 32 aperture, separable Hann; 500 scatterers, 30 dB.

Reduced variants (fewer scanlines/elements/depths, same physics) keep host
oracle runs short in tests.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field, replace

import numpy as np

from . import _native as N
from .geometry import BeamformerConfig, RxGeometry


@dataclass
class Workload:
    name: str
    bf: BeamformerConfig
    fs_hz: float
    num_channels: int
    num_samples: int
    pulse_hz: float
    frac_bw: float = 0.6
    snr_db: float = 30.0
    peak: float = 16384.0
    num_scatterers: int = 200
    box_x: tuple = (-19.0, 19.0)        # mm lateral
    box_depth: tuple = (1.0, 40.0)      # mm
    box_z: tuple = (0.0, 0.0)           # mm elevation
    scatter_seed: int = 0
    noise_seed: int = 1
    beam_sigma_mm: float = 1.0
    # receive beamforming parameters (BeamformingNode XML keys)
    f_number: float = 1.0
    window: int = N.WINDOW_HANN
    window_parameter: float = 0.0
    speed_of_sound: float = 1540.0      # m/s, the node's speedOfSound
    interpolate_transmits: bool = False
    additional_offset: int = 0
    _geom: RxGeometry | None = field(default=None, repr=False, compare=False)

    @property
    def num_tx(self) -> int:
        return self.bf.num_scanlines[0] * self.bf.num_scanlines[1]

    def geometry(self) -> RxGeometry:
        if self._geom is None:
            self._geom = RxGeometry.from_config(self.bf)
        return self._geom

    @property
    def rf_shape(self) -> tuple:
        return (self.num_tx, self.num_channels, self.num_samples)

    @property
    def out_shape(self) -> tuple:
        g = self.geometry()
        return (g.num_depths, g.num_rx)

    def scatterers(self) -> np.ndarray:
        rng = np.random.default_rng(self.scatter_seed)
        n = self.num_scatterers
        sc = np.zeros((n, 4))
        sc[:, 0] = rng.uniform(*self.box_x, n)
        sc[:, 1] = rng.uniform(*self.box_depth, n)
        sc[:, 2] = rng.uniform(*self.box_z, n) if self.box_z[1] > self.box_z[0] else 0.0
        sc[:, 3] = rng.uniform(-1.0, 1.0, n)
        return sc

    def synth_params(self, scatterers: np.ndarray | None = None):
        """SynthParamsC plus the arrays it points into (keep both alive)."""
        g = self.geometry()
        origin, direction, t0, first, last = g.transmit_tables(self.num_tx)
        elem = g.element_positions()
        sc = self.scatterers() if scatterers is None else np.ascontiguousarray(scatterers, dtype=np.float64)
        keep = [np.ascontiguousarray(a) for a in (sc, origin, direction, t0, first, last, elem)]
        sc, origin, direction, t0, first, last, elem = keep
        p = N.SynthParamsC()
        p.num_tx, p.num_channels, p.num_samples = self.num_tx, self.num_channels, self.num_samples
        p.num_elements = g.num_elements
        p.sampling_frequency_hz = self.fs_hz
        p.speed_of_sound_m_per_s = self.bf.speed_of_sound
        p.center_frequency_hz = self.pulse_hz
        p.fractional_bandwidth = self.frac_bw
        p.snr_db = self.snr_db
        p.peak = self.peak
        p.beam_sigma_mm = self.beam_sigma_mm
        p.noise_seed = self.noise_seed
        p.num_scatterers = sc.shape[0]
        p.scatterers = N.ptr(sc, C.c_double)
        p.tx_origin = N.ptr(origin, C.c_double)
        p.tx_direction = N.ptr(direction, C.c_double)
        p.tx_t0 = N.ptr(t0, C.c_double)
        p.tx_first_element = N.ptr(first, C.c_uint32)
        p.tx_last_element = N.ptr(last, C.c_uint32)
        p.element_pos = N.ptr(elem, C.c_double)
        return p, keep

    def rf_host(self, dtype=np.int16, scatterers: np.ndarray | None = None) -> np.ndarray:
        """Synthesize the RF on the host: [tx][channel][sample]."""
        p, keep = self.synth_params(scatterers)
        out = np.zeros(self.rf_shape, dtype=dtype)
        code = N.SUPRA_INT16 if dtype == np.int16 else N.SUPRA_FLOAT32
        N.check(N.lib().supra_synth_rf_host(C.byref(p), out.ctypes.data_as(C.c_void_p), code))
        del keep
        return out

    def rf_device(self, out_ptr: int, dtype=np.int16, stream: int = 0,
                  scatterers: np.ndarray | None = None) -> None:
        """Synthesize the RF into a device buffer (e.g. a torch tensor's data_ptr())."""
        p, keep = self.synth_params(scatterers)
        code = N.SUPRA_INT16 if dtype == np.int16 else N.SUPRA_FLOAT32
        N.check(N.lib().supra_synth_rf_device(C.byref(p), C.c_void_p(out_ptr), code, C.c_void_p(stream)))
        del keep

    def das_kwargs(self) -> dict:
        return dict(f_number=self.f_number, speed_of_sound_mm_per_s=self.speed_of_sound * 1000.0,
                    window_type=self.window, window_parameter=self.window_parameter,
                    interpolate_transmits=self.interpolate_transmits, additional_offset=self.additional_offset)


def linear_2d(**overrides) -> Workload:
    """128-element linear array, 2080 x 128 image (the 2D benchmark)."""
    bf = BeamformerConfig(scan_type="linear", transducer_type="linear", element_layout=(128, 1),
                          pitch=(0.3, 0.3), num_scanlines=(128, 1), max_aperture=(64, 1),
                          tx_max_aperture=(64, 1), depth=40.0, num_depths=2080, tx_focus_active=True,
                          tx_focus_depth=20.0, speed_of_sound=1540.0)
    w = Workload(name="linear_2d", bf=bf, fs_hz=40e6, num_channels=64, num_samples=2080, pulse_hz=7e6,
                 num_scatterers=200, box_x=(-19.0, 19.0), box_depth=(1.0, 40.0), scatter_seed=0, noise_seed=1,
                 beam_sigma_mm=1.0)
    return replace(w, **overrides) if overrides else w


def matrix_3d(**overrides) -> Workload:
    """32x32 matrix probe, 64x64 scanlines x 1000 depths (the 3D benchmark)."""
    sector = 2.0 * math.degrees(math.atan(20.0 / 40.0))
    bf = BeamformerConfig(scan_type="biphased", transducer_type="planar", element_layout=(32, 32),
                          pitch=(0.3, 0.3), num_scanlines=(64, 64), max_aperture=(32, 32),
                          tx_max_aperture=(32, 32), depth=40.0, num_depths=1000, sector_angle_deg=(sector, sector),
                          tx_focus_active=True, tx_focus_depth=20.0, speed_of_sound=1540.0)
    w = Workload(name="matrix_3d", bf=bf, fs_hz=20e6, num_channels=1024, num_samples=1000, pulse_hz=3.5e6,
                 num_scatterers=500, box_x=(-20.0, 20.0), box_depth=(1.0, 40.0), box_z=(-20.0, 20.0),
                 scatter_seed=2, noise_seed=3, beam_sigma_mm=1.5)
    return replace(w, **overrides) if overrides else w


def matrix_3d_small(n_el: int = 8, n_lines: int = 8, n_depths: int = 200, n_samples: int = 1100,
                    **overrides) -> Workload:
    """Reduced matrix probe with the same physics (host-oracle sized)."""
    sector = 2.0 * math.degrees(math.atan(20.0 / 40.0))
    bf = BeamformerConfig(scan_type="biphased", transducer_type="planar", element_layout=(n_el, n_el),
                          pitch=(0.3, 0.3), num_scanlines=(n_lines, n_lines), max_aperture=(n_el, n_el),
                          tx_max_aperture=(n_el, n_el), depth=40.0, num_depths=n_depths,
                          sector_angle_deg=(sector, sector), tx_focus_active=True, tx_focus_depth=20.0,
                          speed_of_sound=1540.0)
    w = Workload(name=f"matrix_3d_{n_el}x{n_el}_{n_lines}x{n_lines}x{n_depths}", bf=bf, fs_hz=20e6,
                 num_channels=n_el * n_el, num_samples=n_samples, pulse_hz=3.5e6, num_scatterers=60,
                 box_x=(-20.0, 20.0), box_depth=(1.0, 40.0), box_z=(-20.0, 20.0), scatter_seed=2,
                 noise_seed=3, beam_sigma_mm=1.5)
    return replace(w, **overrides) if overrides else w


def linear_acquisition(num_rx: int, num_depths: int, aperture: int, fs_hz: float, num_scatterers: int,
                       scatter_seed: int, noise_seed: int, **overrides) -> Workload:
    """A linear-array acquisition of the uneven stream: num_rx elements at
    0.3 mm pitch and as many focused transmits (one scanline each), a receive
    and transmit aperture of `aperture` elements (at most the array) on
    `aperture` channels, and num_depths RF samples imaged to the depth they
    reach (c = 1540 m/s)."""
    c_mm_per_s = 1540.0e3
    elements = min(aperture, num_rx)
    depth = num_depths / fs_hz * c_mm_per_s / 2.0
    half = num_rx * 0.3 / 2.0
    bf = BeamformerConfig(scan_type="linear", transducer_type="linear", element_layout=(num_rx, 1),
                          pitch=(0.3, 0.3), num_scanlines=(num_rx, 1), max_aperture=(elements, 1),
                          tx_max_aperture=(elements, 1), depth=depth, num_depths=num_depths, tx_focus_active=True,
                          tx_focus_depth=depth / 2.0, speed_of_sound=1540.0)
    w = Workload(name=f"linear_{num_rx}x{num_depths}_a{aperture}_{fs_hz / 1e6:.0f}MHz", bf=bf, fs_hz=fs_hz,
                 num_channels=aperture, num_samples=num_depths, pulse_hz=fs_hz / 6.0,
                 num_scatterers=num_scatterers, box_x=(-half, half), box_depth=(1.0, depth),
                 scatter_seed=scatter_seed, noise_seed=noise_seed, beam_sigma_mm=1.0)
    return replace(w, **overrides) if overrides else w


def uneven_stream(num: int = 48, seed: int = 5):
    """The uneven batched stream: `num` seeded acquisitions with scanlines in
    {64, 128, 192, 256}, depth samples in [1024, 4096], aperture in
    {32, 64, 128} channels, fs in {20, 40} MHz and 100-400 scatterers, each
    beamformed by delay-and-sum or by minimum variance (L = aperture / 2,
    T = 5).  Returns [(Workload, mode)] with mode "das" or "mv"."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(num):
        n_rx = int(rng.choice([64, 128, 192, 256]))
        n_d = int(rng.integers(1024, 4097))
        aperture = int(rng.choice([32, 64, 128]))
        fs = float(rng.choice([20e6, 40e6]))
        n_sc = int(rng.integers(100, 401))
        mode = "mv" if rng.random() < 0.5 else "das"
        out.append((linear_acquisition(n_rx, n_d, aperture, fs, n_sc, scatter_seed=1000 + i,
                                       noise_seed=2000 + i), mode))
    return out


WORKLOADS = {"linear_2d": linear_2d, "matrix_3d": matrix_3d}
