"""Pins the host oracle's reverse mode of the 2D DAS (oracle_das_adjoint):
the dot-product identity <DAS x, y> = <x, DAS^T y> against the oracle's own
forward (oracle_das, pinned in test_oracle_das.py), dL/dc against FP64 central
differences of that forward, and a pure-Python loop on a tiny array.

Central differences in c see the kinks of linear interpolation: dy/dtau jumps
where tau crosses an integer, so a difference over +-delta measures a secant
wherever a kink lies inside the interval.  The check at delta_c = 0.5 m/s
therefore uses RF that is linear in the sample index (no kinks: the
interpolant is exact), and the check on the real workload RF uses a small
delta (the fraction of straddled kinks, hence the error, scales with delta).
Pixels whose aperture membership, window or sample range changes inside the
interval (rounding-level ties; they show as changes of the image of a
constant and of a channel-ramp RF) are left out of the loss."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from supra_amd import _native as N
from supra_amd import workloads as W
from supra_amd.geometry import BeamformerConfig

FLT_MAX = 3.4028234663852886e38


def fwd64(w, g, rf, c_mm, **kw):
    args = dict(w.das_kwargs())
    args.update(kw)
    args["speed_of_sound_mm_per_s"] = c_mm
    return O.das(g, rf, w.fs_hz, precision=64, **args)


def adj64(w, g, rf, c_mm, grad_image, **kw):
    args = dict(w.das_kwargs())
    args.update(kw)
    args.pop("speed_of_sound_mm_per_s")
    return O.das_adjoint(g, rf, w.fs_hz, args.pop("f_number"), c_mm, args.pop("window_type"), grad_image,
                         **args)


def smooth_mask(w, g, c_mm, dc, **kw):
    """Pixels whose membership, window weights and sample range do not change on [c-dc, c+dc]."""
    shape = w.rf_shape
    ones = np.ones(shape, np.float32)
    ramp = np.broadcast_to(np.sqrt(np.arange(shape[1]) + 2.0)[None, :, None], shape).astype(np.float32)
    m = np.ones((g.num_depths, g.num_rx), bool)
    for sig in (ones, np.ascontiguousarray(ramp)):
        ims = [fwd64(w, g, sig, c, **kw) for c in (c_mm - dc, c_mm, c_mm + dc)]
        m &= (ims[0] == ims[1]) & (ims[1] == ims[2]) & (np.abs(ims[1]) < FLT_MAX)
    return m


CASES = [
    ("linear_hann", dict(), dict()),
    ("hamming_offset", dict(window=N.WINDOW_HAMMING, additional_offset=7), dict()),
    ("subdivided_interp", dict(interpolate_transmits=True),
     dict(num_scanlines=(64, 1), rx_scanline_subdivision=(2, 1))),
]


def workload(overrides, bf_overrides):
    w = W.linear_2d(num_scatterers=60, **overrides)
    if bf_overrides:
        from dataclasses import replace
        w = replace(w, bf=replace(w.bf, **bf_overrides))
    return w


@pytest.mark.parametrize("name,ov,bfov", CASES, ids=[c[0] for c in CASES])
def test_dot_product_identity(name, ov, bfov):
    w = workload(ov, bfov)
    g = w.geometry()
    rng = np.random.default_rng(7)
    x = rng.normal(size=w.rf_shape).astype(np.float32)
    c = 1510e3
    y = fwd64(w, g, x, c)
    v = rng.normal(size=y.shape)
    v[~(np.abs(y) < FLT_MAX)] = 0.0
    grad_rf, _, _ = adj64(w, g, x, c, v)
    lhs = float((np.where(np.abs(y) < FLT_MAX, y, 0.0) * v).sum())
    rhs = float((x.astype(np.float64) * grad_rf[0]).sum())
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)
    # and the adjoint of a one-hot image is the pixel's RF weights: its sum
    # against a constant RF is the forward of that constant.
    assert np.isfinite(grad_rf).all()


@pytest.mark.parametrize("c_mm", [1450e3, 1540e3, 1600e3])
def test_dc_central_difference_linear_rf(c_mm):
    """delta_c = 0.5 m/s on RF linear in the sample index: exact interpolation, smooth in c."""
    w = W.linear_2d(num_scatterers=10)
    g = w.geometry()
    rng = np.random.default_rng(int(c_mm))
    n_tx, n_ch, n_s = w.rf_shape
    # Positive slopes and positive pixel weights: every term of dL/dc has the
    # same sign, so the check is not blurred by cancellation.
    rf = (np.arange(n_s)[None, None, :] * rng.uniform(0.5, 1.5, (n_tx, n_ch, 1)) +
          rng.uniform(-100, 100, (n_tx, n_ch, 1))).astype(np.float32)
    dc = 500.0
    v = np.where(smooth_mask(w, g, c_mm, dc), rng.uniform(0.5, 1.5, size=(g.num_depths, g.num_rx)), 0.0)
    f = lambda c: float((fwd64(w, g, rf, c) * v).sum())
    fd = (f(c_mm + dc) - f(c_mm - dc)) / (2 * dc)
    _, gc, _ = adj64(w, g, rf, c_mm, v, want_rf=False)
    assert abs(gc[0] - fd) <= 1e-6 * abs(fd), (gc[0], fd)


@pytest.mark.parametrize("c_mm,dc,tol", [(1450e3, 0.1, 1e-4), (1600e3, 0.1, 1e-4), (1540e3, 1.0, 5e-4)])
def test_dc_central_difference_workload_rf(c_mm, dc, tol):
    """Small delta on the scatterer RF (kinks inside +-delta are rare).  The
    functional weights each pixel by the sign of its own difference quotient,
    so the per-pixel derivatives add without cancelling."""
    w = W.linear_2d()
    g = w.geometry()
    rf = w.rf_host(dtype=np.float32)
    m = smooth_mask(w, g, c_mm, dc)
    fd_pixels = (fwd64(w, g, rf, c_mm + dc) - fwd64(w, g, rf, c_mm - dc)) / (2 * dc)
    v = np.where(m, np.sign(fd_pixels), 0.0)
    fd = float((v * fd_pixels).sum())
    _, gc, _ = adj64(w, g, rf, c_mm, v, want_rf=False)
    assert abs(gc[0] - fd) <= tol * abs(fd), (gc[0], fd)


def test_rectangular_window_norm_is_one():
    w = W.linear_2d(num_scatterers=10, window=N.WINDOW_RECTANGULAR)
    g = w.geometry()
    rf = w.rf_host(dtype=np.float32)
    _, _, norm = adj64(w, g, rf, 1540e3, np.zeros((g.num_depths, g.num_rx)), want_rf=False, want_norm=True)
    assert set(np.unique(norm)) <= {0.0, 1.0}
    assert (norm == 1.0).mean() > 0.99


@np.errstate(invalid="ignore")
def python_adjoint_2d(g, rf, fs, F, c_mm, lut, grad_image, offset=0):
    """Pure-Python float64 reverse mode of sampleBeamform2D (one transmit per scanline)."""
    dt = 1.0 / fs
    fac = 1.0 / (c_mm * dt)
    n_tx, n_ch, n_s = rf.shape
    grad = np.zeros_like(rf, dtype=np.float64)
    dc = 0.0
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
            terms = []
            wsum = 0.0
            for e in range(tp["first"][0], tp["last"][0]):
                xe = float(g.element_x[e]) * fac
                if abs(xe - x) <= aDT:
                    rel = (xe - x) * inv
                    rel = -1.0 if math.isnan(rel) else min(max(rel, -1.0), 1.0)
                    wgt = lut[int(math.floor(31.5 * (rel + 1.0) + 0.5))]
                    wsum += wgt
                    geo = math.sqrt((xe - (x + dirx * d)) ** 2 + (diry * d) ** 2) + d
                    terms.append((e % n_ch, wgt, t0 + geo + offset, geo))
            if not terms or wsum == 0:
                continue
            beta = grad_image[r, s] * len(terms) / wsum
            for ch, wgt, tau, geo in terms:
                i = math.floor(tau)
                f = tau - i
                line = rf[tp["tx_idx"], ch]
                if 0 <= i < n_s - 1:
                    grad[tp["tx_idx"], ch, i] += beta * wgt * (1 - f)
                    grad[tp["tx_idx"], ch, i + 1] += beta * wgt * f
                    dc += beta * wgt * (line[i + 1] - line[i]) * geo
                elif 0 <= i < n_s and f == 0:
                    grad[tp["tx_idx"], ch, i] += beta * wgt
    return grad, -dc / c_mm


def test_brute_force_tiny_array():
    bf = BeamformerConfig(num_scanlines=(8, 1), element_layout=(8, 1), max_aperture=(6, 1), tx_max_aperture=(6, 1),
                          num_depths=40, depth=2.0, tx_focus_depth=1.0)
    w = W.linear_2d(bf=bf, num_channels=8, num_samples=64, num_scatterers=5, box_x=(-1.0, 1.0),
                    box_depth=(0.2, 2.0), window=N.WINDOW_HAMMING, f_number=0.7)
    g = w.geometry()
    rf = w.rf_host(dtype=np.float32)
    lut = O.window_lut(N.WINDOW_HAMMING, 0.0, 64)
    rng = np.random.default_rng(3)
    for off in (0, 5):
        gi = rng.normal(size=(g.num_depths, g.num_rx))
        ref_rf, ref_c = python_adjoint_2d(g, rf.astype(np.float64), w.fs_hz, w.f_number, 1530e3, lut, gi, off)
        got_rf, got_c, _ = adj64(w, g, rf, 1530e3, gi, additional_offset=off)
        assert np.allclose(got_rf[0], ref_rf, rtol=1e-10, atol=1e-10)
        assert abs(got_c[0] - ref_c) <= 1e-10 * max(1.0, abs(ref_c))


def test_fp32_adjoint_agrees_with_fp64():
    """Single precision differs from double where a membership or window
    index flips and where W is a sum of tiny edge weights (shallow pixels):
    normwise agreement, as for the forward."""
    w = W.linear_2d(num_scatterers=50)
    g = w.geometry()
    rf = w.rf_host(dtype=np.float32)
    rng = np.random.default_rng(11)
    gi = rng.normal(size=(g.num_depths, g.num_rx))
    r64, c64, _ = adj64(w, g, rf, 1490e3, gi, precision=64)
    r32, c32, _ = adj64(w, g, rf, 1490e3, gi, precision=32)
    assert np.linalg.norm(r32 - r64) <= 1e-2 * np.linalg.norm(r64)
    assert np.quantile(np.abs(r32 - r64), 0.99) <= 1e-3 * np.abs(r64).max()
    assert abs(c32[0] - c64[0]) <= 5e-3 * abs(c64[0])


def test_mixed_precision_mode():
    """precision=6432 (FP64 values on FP32 decisions): identical to FP64 on a
    tiny array where float and double make the same decisions, and within
    single-precision rounding of the FP32 mode on the workload."""
    bf = BeamformerConfig(num_scanlines=(8, 1), element_layout=(8, 1), max_aperture=(6, 1), tx_max_aperture=(6, 1),
                          num_depths=40, depth=2.0, tx_focus_depth=1.0)
    w = W.linear_2d(bf=bf, num_channels=8, num_samples=64, num_scatterers=5, box_x=(-1.0, 1.0),
                    box_depth=(0.2, 2.0), window=N.WINDOW_HAMMING, f_number=0.7)
    g = w.geometry()
    rf = w.rf_host(dtype=np.float32)
    gi = np.random.default_rng(4).normal(size=(g.num_depths, g.num_rx))
    _, _, n64 = adj64(w, g, rf, 1530e3, gi, precision=64, want_norm=True)
    _, _, nmx = adj64(w, g, rf, 1530e3, gi, precision=6432, want_norm=True)
    # pixels where float and double agree on membership and window indices
    same = (np.abs(nmx - n64) <= 1e-12 * np.abs(n64))[0, 0].T
    assert 0.8 < same.mean() < 1.0        # the tiny array has element-aligned ties
    gi = np.where(same, gi, 0.0)
    r64, c64, _ = adj64(w, g, rf, 1530e3, gi, precision=64)
    rmx, cmx, _ = adj64(w, g, rf, 1530e3, gi, precision=6432)
    assert np.allclose(rmx, r64, rtol=1e-12, atol=1e-12) and abs(cmx[0] - c64[0]) <= 1e-12 * abs(c64[0])

    w = W.linear_2d(num_scatterers=50)
    g = w.geometry()
    rf = w.rf_host(dtype=np.float32)
    gi = np.random.default_rng(12).normal(size=(g.num_depths, g.num_rx))
    r32, c32, _ = adj64(w, g, rf, 1520e3, gi, precision=32)
    rmx, cmx, _ = adj64(w, g, rf, 1520e3, gi, precision=6432)
    assert np.linalg.norm(r32 - rmx) <= 1e-3 * np.linalg.norm(rmx)
    assert abs(c32[0] - cmx[0]) <= 1e-4 * abs(cmx[0])


def test_decision_diff_classifier():
    """das_decision_diff marks the pixels where FP32 and FP64 take different
    discrete decisions: a rectangular window never sets the window flag, and
    outside the marked pixels the FP32 and FP64 images agree to
    single-precision rounding of the delays, while the marked pixels hold the
    large differences."""
    w = W.linear_2d(num_scatterers=60)
    g = w.geometry()
    rf = w.rf_host()
    kw = w.das_kwargs()
    rect = O.das_decision_diff(g, w.num_samples, w.fs_hz, 1.0, 1540e3, N.WINDOW_RECTANGULAR)
    assert not (rect & 2).any()
    fl = O.das_decision_diff(g, w.num_samples, w.fs_hz, 1.0, 1540e3, N.WINDOW_HANN)
    assert (fl & 2).any()
    y32 = O.das(g, rf, w.fs_hz, **kw).astype(np.float64)
    y64 = O.das(g, rf, w.fs_hz, precision=64, **kw)
    ok = (np.abs(y32) < 1e30) & (np.abs(y64) < 1e30)
    scale = np.abs(y64[ok]).max()
    clean = ok & (fl == 0)
    tied = ok & (fl != 0)
    assert np.abs(y32[clean] - y64[clean]).max() <= 1e-4 * scale
    assert np.abs(y32[tied] - y64[tied]).max() >= 1e-3 * scale
    same = O.das_decision_diff(g, w.num_samples, w.fs_hz, 1.0, 1540e3, N.WINDOW_HANN, precision_b=32)
    assert not same.any()
    # across speeds of sound only the delay decisions (floor, range) move in
    # exact arithmetic; membership and window changes are rounding ties
    moved = O.das_decision_diff(g, w.num_samples, w.fs_hz, 1.0, 1540e3, N.WINDOW_HANN, precision_a=64,
                                speed_of_sound_b_mm_per_s=1545e3)
    assert (moved & 4).mean() > 0.5 and (moved & 3).mean() < 0.2
