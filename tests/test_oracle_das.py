"""Pins the host oracle's delay-and-sum (RxSampleBeamformerDelayAndSum.h,
RxBeamformerCommon.h, RxBeamformerCuda.cu kernels) with properties whose
expected values come from the millimetre geometry, not from the oracle's code:
constant and ramp RF, linearity, point-scatterer focus, the TestSignal known
answer, and a brute-force loop on a tiny array."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from supra_amd import _native as N
from supra_amd import workloads as W
from supra_amd.geometry import BeamformerConfig, RxGeometry


def pixel_points(g):
    """World position of every pixel [depth, scanline, xyz] in mm: position + depth * direction."""
    pos = g.scanlines["position"]            # (x, depth axis, elevation)
    dirs = g.scanlines["direction"]
    d = g.depths.astype(np.float64)
    return pos[None, :, :] + d[:, None, None] * dirs[None, :, :]


def members_and_delays(w, g, f_number, k=0):
    """Aperture membership and two-way delays [samples] per (depth, scanline, element),
    from the mm geometry: transmit path = depth along the scanline, receive path =
    pixel -> element distance, plus the scanline's initial delay."""
    c = w.speed_of_sound * 1000.0
    fs = w.fs_hz
    P = pixel_points(g)
    ex = g.element_x.astype(np.float64)
    ey = g.element_y.astype(np.float64)
    d = g.depths.astype(np.float64)
    s = g.scanlines
    nd, nr = P.shape[:2]
    E = np.stack([ex, np.zeros_like(ex), ey], 1)           # elements at depth 0
    first = s["tx"][:, k]["first"]
    last = s["tx"][:, k]["last"]
    t0 = s["tx"][:, k]["initial_delay"] * fs
    if not g.is3d:
        lat = np.abs(ex[None, :] - s["position"][:, 0:1])  # [nr, ne]
        inap = np.zeros((nr, g.num_elements), bool)
        for i in range(nr):
            inap[i, first[i, 0]:last[i, 0]] = True
        member = (lat[None, :, :] <= d[:, None, None] / (2 * f_number)) & inap[None]
    else:
        r2 = (ex[None, :] - s["position"][:, 0:1]) ** 2 + (ey[None, :] - s["position"][:, 2:3]) ** 2
        lx = g.element_layout[0]
        inap = np.zeros((nr, g.num_elements), bool)
        for i in range(nr):
            for exi in range(first[i, 0], last[i, 0]):
                for eyi in range(first[i, 1], last[i, 1]):
                    inap[i, exi + eyi * lx] = True
        member = (r2[None] <= (d[:, None, None] / (2 * f_number)) ** 2) & inap[None]
    rx = np.linalg.norm(P[:, :, None, :] - E[None, None, :, :], axis=-1)   # [nd, nr, ne]
    tau = t0[None, :, None] + (rx + d[:, None, None]) / c * fs
    return member, tau


def boundary_safe(member, g, f_number, rel=1e-5):
    """Pixels whose membership decision is not within `rel` of its threshold."""
    d = g.depths.astype(np.float64)
    s = g.scanlines
    if not g.is3d:
        lat = np.abs(g.element_x[None, :].astype(np.float64) - s["position"][:, 0:1])
        margin = np.abs(lat[None] - d[:, None, None] / (2 * f_number))
    else:
        r = np.sqrt((g.element_x[None, :] - s["position"][:, 0:1]) ** 2 +
                    (g.element_y[None, :] - s["position"][:, 2:3]) ** 2)
        margin = np.abs(r[None] - d[:, None, None] / (2 * f_number))
    return (margin > rel * np.maximum(d[:, None, None], 1.0)).all(-1)


def das64(w, g, rf, **kw):
    args = dict(w.das_kwargs())
    args.update(kw)
    return O.das(g, rf, w.fs_hz, precision=64, **args)


def test_constant_rf_gives_k_times_n():
    w = W.linear_2d()
    g = w.geometry()
    k = 123
    rf = np.full(w.rf_shape, k, dtype=np.int16)
    out = das64(w, g, rf)
    member, tau = members_and_delays(w, g, w.f_number)
    n = member.sum(-1)
    ok = boundary_safe(member, g, w.f_number) & (np.where(member, tau, 0).max(-1) < w.num_samples - 2) & (n > 1)
    assert ok.sum() > 100000
    assert np.allclose(out[ok], k * n[ok], rtol=1e-12)


@pytest.mark.parametrize("steer", [0.0, 9.0])
def test_ramp_rf_gives_sum_of_delays(steer):
    bf = BeamformerConfig(num_scanlines=(48, 1), element_layout=(96, 1), max_aperture=(48, 1),
                          tx_max_aperture=(48, 1), num_depths=600, depth=30.0,
                          steering_angle_deg=(steer, 0.0))
    w = W.linear_2d(bf=bf, num_channels=96, num_samples=2000, window=N.WINDOW_RECTANGULAR, f_number=1.3)
    g = w.geometry()
    a, b = 3.0, 0.25
    rf = np.broadcast_to((a + b * np.arange(w.num_samples)).astype(np.float32), w.rf_shape).copy()
    out = das64(w, g, rf)
    member, tau = members_and_delays(w, g, w.f_number)
    expected = np.where(member, a + b * tau, 0.0).sum(-1)
    ok = boundary_safe(member, g, w.f_number) & (np.where(member, tau, 0).max(-1) < w.num_samples - 2)
    ok &= member.sum(-1) > 0
    assert ok.sum() > 10000
    err = np.abs(out[ok] - expected[ok]) / np.abs(expected[ok])
    assert err.max() < 1e-9, err.max()


def test_linearity():
    w = W.linear_2d(num_scatterers=40)
    g = w.geometry()
    r1 = w.rf_host(dtype=np.float32)
    r2 = W.linear_2d(num_scatterers=30, scatter_seed=5, noise_seed=6).rf_host(dtype=np.float32)
    lhs = das64(w, g, (0.5 * r1 - 2.0 * r2).astype(np.float32))
    rhs = 0.5 * das64(w, g, r1) - 2.0 * das64(w, g, r2)
    fin = np.isfinite(lhs) & (np.abs(lhs) < 1e30)
    assert np.abs(lhs[fin] - rhs[fin]).max() <= 1e-6 * np.abs(rhs[fin]).max()


@pytest.mark.parametrize("x0,z0", [(0.15, 22.0), (-7.35, 13.0), (9.75, 31.5)])
def test_point_scatterer_focuses_on_its_pixel(x0, z0):
    w = W.linear_2d(snr_db=200.0, window=N.WINDOW_RECTANGULAR)
    g = w.geometry()
    sc = np.array([[x0, z0, 0.0, 1.0]])
    rf = w.rf_host(dtype=np.float32, scatterers=sc)
    out = np.abs(das64(w, g, rf))
    out[:5] = 0
    r, s = np.unravel_index(np.argmax(out), out.shape)
    xs = g.scanlines["position"][:, 0]
    assert abs(xs[s] - x0) < 0.3 + 1e-9            # within one scanline pitch
    assert abs(g.depths[r] - z0) < 0.06            # within a quarter wavelength (0.22 mm)


def test_test_signal_known_answer():
    w = W.linear_2d()
    g = w.geometry()
    rf = np.zeros(w.rf_shape, dtype=np.int16)
    out = das64(w, g, rf, test_signal=True)
    P = pixel_points(g)
    dist = np.min(np.stack([np.hypot(P[..., 0] - 6.0 * k, P[..., 1] - 30.0) for k in range(-3, 4)]), axis=0)
    expected = np.where(dist <= 2.0, 1000.0, 0.0)
    safe = np.abs(dist - 2.0) > 1e-4
    assert np.array_equal(out[safe], expected[safe])
    assert (expected == 1000).sum() > 1000


def test_fp32_agrees_with_fp64():
    w = W.linear_2d()
    g = w.geometry()
    rf = w.rf_host()
    o32 = O.das(g, rf, w.fs_hz, **w.das_kwargs()).astype(np.float64)
    o64 = das64(w, g, rf)
    fin = np.isfinite(o64) & (np.abs(o64) < 1e30) & (np.abs(o32) < 1e30)
    den = np.abs(o64[fin]).max()
    err = np.abs(o32[fin] - o64[fin]) / den
    # The few largest differences are pixels where single-precision geometry
    # flips an element across the aperture edge or a window-table index.
    assert err.max() < 5e-3, err.max()
    assert np.quantile(err, 0.99) < 2e-4, np.quantile(err, 0.99)


def python_das_2d(g, rf, fs, F, c_mm, lut, offset=0):
    """Pure-Python float64 loop of sampleBeamform2D for the brute-force check."""
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
            acc = wsum = 0.0
            n = 0
            for e in range(tp["first"][0], tp["last"][0]):
                xe = float(g.element_x[e]) * fac
                if abs(xe - x) <= aDT:
                    rel = (xe - x) * inv
                    rel = -1.0 if math.isnan(rel) else min(max(rel, -1.0), 1.0)
                    wgt = lut[int(math.floor(31.5 * (rel + 1.0) + 0.5))]
                    wsum += wgt
                    n += 1
                    tau = t0 + math.sqrt((xe - (x + dirx * d)) ** 2 + (diry * d) ** 2) + d + offset
                    i = math.floor(tau)
                    f = tau - i
                    ch = e % n_ch
                    if 0 <= i < n_s - 1:
                        acc += wgt * ((1 - f) * rf[tp["tx_idx"], ch, i] + f * rf[tp["tx_idx"], ch, i + 1])
                    elif 0 <= i < n_s and f == 0:
                        acc += wgt * rf[tp["tx_idx"], ch, i]
            out[r, s] = (acc / wsum * n if wsum != 0 else math.nan) if n > 0 else 0.0
    return out


def test_brute_force_tiny_array():
    bf = BeamformerConfig(num_scanlines=(8, 1), element_layout=(8, 1), max_aperture=(6, 1), tx_max_aperture=(6, 1),
                          num_depths=40, depth=2.0, tx_focus_depth=1.0)
    w = W.linear_2d(bf=bf, num_channels=8, num_samples=64, num_scatterers=5, box_x=(-1.0, 1.0),
                    box_depth=(0.2, 2.0), window=N.WINDOW_HAMMING, f_number=0.7)
    g = w.geometry()
    rf = w.rf_host(dtype=np.float32)
    lut = O.window_lut(N.WINDOW_HAMMING, 0.0, 64)
    for off in (0, 5):
        ref = python_das_2d(g, rf.astype(np.float64), w.fs_hz, w.f_number, 1540e3, lut, offset=off)
        got = das64(w, g, rf, additional_offset=off)
        m = np.isfinite(ref) & np.isfinite(got) & (np.abs(got) < 1e30)
        assert np.allclose(got[m], ref[m], rtol=1e-10, atol=1e-9)


def test_3d_constant_and_ramp():
    w = W.matrix_3d_small(n_el=10, n_lines=6, n_depths=250, window=N.WINDOW_RECTANGULAR)
    g = w.geometry()
    member, tau = members_and_delays(w, g, w.f_number)
    safe = boundary_safe(member, g, w.f_number) & (np.where(member, tau, 0).max(-1) < w.num_samples - 2)
    safe &= member.sum(-1) > 0
    assert safe.sum() > 3000
    rf = np.full(w.rf_shape, -77, dtype=np.int16)
    out = das64(w, g, rf)
    assert np.allclose(out[safe], -77 * member.sum(-1)[safe], rtol=1e-12)
    a, b = -2.0, 0.5
    rf = np.broadcast_to((a + b * np.arange(w.num_samples)).astype(np.float32), w.rf_shape).copy()
    out = das64(w, g, rf)
    expected = np.where(member, a + b * tau, 0.0).sum(-1)
    assert np.allclose(out[safe], expected[safe], rtol=1e-9)


def test_negative_delays_2d_skipped_3d_extrapolated():
    """additionalOffset pushes shallow delays below zero: 2D drops those samples
    (int32 floor, unsigned compare), 3D reads sample 0 with a negative fraction
    (saturating uint32 floor), i.e. extrapolates the ramp."""
    a, b = 5.0, 1.0
    w2 = W.linear_2d(window=N.WINDOW_RECTANGULAR, additional_offset=-400)
    g2 = w2.geometry()
    rf2 = np.broadcast_to((a + b * np.arange(w2.num_samples)).astype(np.float32), w2.rf_shape).copy()
    out2 = das64(w2, g2, rf2)
    m2, t2 = members_and_delays(w2, g2, w2.f_number)
    t2 = t2 - 400
    exp2 = np.where(m2 & (t2 >= 0), a + b * t2, 0.0).sum(-1)
    ok2 = boundary_safe(m2, g2, w2.f_number) & (np.abs(t2 - np.round(t2)) > 1e-6).all(-1) & (m2.sum(-1) > 0)
    ok2 &= np.where(m2, t2, 0).max(-1) < w2.num_samples - 2
    assert ok2.sum() > 1000 and (np.where(m2, t2, 1).min(-1)[ok2] < 0).sum() > 100
    assert np.allclose(out2[ok2], exp2[ok2], rtol=1e-9)

    w3 = W.matrix_3d_small(n_el=8, n_lines=6, n_depths=250, window=N.WINDOW_RECTANGULAR, additional_offset=-60)
    g3 = w3.geometry()
    rf3 = np.broadcast_to((a + b * np.arange(w3.num_samples)).astype(np.float32), w3.rf_shape).copy()
    out3 = das64(w3, g3, rf3)
    m3, t3 = members_and_delays(w3, g3, w3.f_number)
    t3 = t3 - 60
    exp3 = np.where(m3, a + b * t3, 0.0).sum(-1)
    ok3 = boundary_safe(m3, g3, w3.f_number) & (m3.sum(-1) > 0) & (np.where(m3, t3, 0).max(-1) < w3.num_samples - 2)
    assert (np.where(m3, t3, 1).min(-1)[ok3] < 0).sum() > 100
    assert np.allclose(out3[ok3], exp3[ok3], rtol=1e-9)
