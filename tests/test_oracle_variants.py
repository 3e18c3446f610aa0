"""Pins the host oracle's DelayAndStdDev and CoherenceFactorDelayAndSum
beamformers (RxSampleBeamformerDelayAndStdDev.h,
RxSampleBeamformerCoherenceFactorDelayAndSum.h): a pure-Python float64 loop
on a tiny array, and closed forms on constant RF — the coherence factor of a
constant is 1 (CF-DAS = DAS = k n), the standard deviation of a constant
under a rectangular window is sqrt(n) |k| (n - 1) (the reference's `mean`
is the DAS value k n)."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from supra_amd import _native as N
from supra_amd import workloads as W
from supra_amd.geometry import BeamformerConfig

STDDEV, CF = 1, 2


def python_variant_2d(g, rf, fs, F, c_mm, lut, kind, offset=0):
    dt = 1.0 / fs
    fac = 1.0 / (c_mm * dt)
    n_tx, n_ch, n_s = rf.shape
    out = np.zeros((g.num_depths, g.num_rx))
    for s in range(g.num_rx):
        sc = g.scanlines[s]
        x = sc["position"][0] * fac
        dirx, diry = sc["direction"][0], sc["direction"][1]
        med = sc["max_element_distance"][0] * fac
        tp = sc["tx"][0]
        t0 = tp["initial_delay"] / dt
        for r in range(g.num_depths):
            d = float(g.depths[r]) * fac
            aDT = d * dt * c_mm / (2 * F) / c_mm / dt
            m = min(aDT, med)
            inv = math.inf if m == 0 else 1.0 / m
            taps = []
            for e in range(tp["first"][0], tp["last"][0]):
                xe = float(g.element_x[e]) * fac
                if abs(xe - x) <= aDT:
                    rel = (xe - x) * inv
                    rel = -1.0 if math.isnan(rel) else min(max(rel, -1.0), 1.0)
                    wgt = lut[int(math.floor(31.5 * (rel + 1.0) + 0.5))]
                    tau = t0 + math.sqrt((xe - (x + dirx * d)) ** 2 + (diry * d) ** 2) + d + offset
                    i = math.floor(tau)
                    f = tau - i
                    ch = e % n_ch
                    v = 0.0
                    if 0 <= i < n_s - 1:
                        v = (1 - f) * rf[tp["tx_idx"], ch, i] + f * rf[tp["tx_idx"], ch, i + 1]
                    elif 0 <= i < n_s and f == 0:
                        v = rf[tp["tx_idx"], ch, i]
                    taps.append((wgt, v))
            n = len(taps)
            if n == 0:
                continue
            wsum = sum(t[0] for t in taps)
            das = sum(wg * v for wg, v in taps) / wsum * n
            if kind == STDDEV:
                out[r, s] = math.sqrt(sum((wg * v - das) ** 2 for wg, v in taps)) / wsum * n
            else:
                cs = sum(v for _, v in taps)
                te = sum(v * v for _, v in taps)
                out[r, s] = cs * cs / te / n * das if te else math.nan
    return out


@pytest.mark.parametrize("kind", [STDDEV, CF])
def test_variants_brute_force_tiny_array(kind):
    bf = BeamformerConfig(num_scanlines=(8, 1), element_layout=(8, 1), max_aperture=(6, 1), tx_max_aperture=(6, 1),
                          num_depths=40, depth=2.0, tx_focus_depth=1.0)
    w = W.linear_2d(bf=bf, num_channels=8, num_samples=64, num_scatterers=5, box_x=(-1.0, 1.0),
                    box_depth=(0.2, 2.0), window=N.WINDOW_HAMMING, f_number=0.7)
    g = w.geometry()
    rf = w.rf_host(dtype=np.float32)
    lut = O.window_lut(N.WINDOW_HAMMING, 0.0, 64)
    for off in (0, 5):
        ref = python_variant_2d(g, rf.astype(np.float64), w.fs_hz, w.f_number, 1540e3, lut, kind, offset=off)
        args = dict(w.das_kwargs(), additional_offset=off)
        got = O.das(g, rf, w.fs_hz, precision=64, sample_beamformer=kind, **args)
        m = np.isfinite(ref) & np.isfinite(got) & (np.abs(got) < 1e30)
        assert m.sum() > 100
        assert np.allclose(got[m], ref[m], rtol=1e-10, atol=1e-9)


@pytest.mark.parametrize("geom", ["2d", "3d"])
def test_variants_constant_rf_closed_forms(geom):
    if geom == "2d":
        w = W.linear_2d(num_scatterers=1, window=N.WINDOW_RECTANGULAR)
    else:
        w = W.matrix_3d_small(n_el=8, n_lines=5, n_depths=160, n_samples=900, window=N.WINDOW_RECTANGULAR)
    g = w.geometry()
    k = 37.0
    rf = np.full(w.rf_shape, k, dtype=np.float32)
    kw = w.das_kwargs()
    das = O.das(g, rf, w.fs_hz, precision=64, **kw)
    cf = O.das(g, rf, w.fs_hz, precision=64, sample_beamformer=CF, **kw)
    sd = O.das(g, rf, w.fs_hz, precision=64, sample_beamformer=STDDEV, **kw)
    n = das / k
    # depths 10 % .. 70 % of the image, where every member's delay is inside
    # the RF (then DAS = k n with n the member count, a whole number)
    ok = (np.abs(n - np.round(n)) < 1e-9) & (n >= 2)
    ok[: g.num_depths // 10] = False
    ok[int(0.7 * g.num_depths):] = False
    assert ok.sum() > 1000
    assert np.allclose(cf[ok], das[ok], rtol=1e-12)
    assert np.allclose(sd[ok], np.sqrt(n[ok]) * k * (n[ok] - 1), rtol=1e-12)
