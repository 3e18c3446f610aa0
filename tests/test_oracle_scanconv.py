"""The scan-conversion oracle pinned by what the geometry fixes: closed forms
of parallel (linear array) and common-origin (phased array) scanlines in 2D
and 3D, the defining identity pixel = base(t) + d * direction(t) on curved
arrays, the scanned area, and the kernel's bilinear interpolation of data
that is linear in position.  (The reference ships no scan-conversion tests
or vectors.)"""
import numpy as np
import pytest

from oracle import scanconv as SC


def linear_2d(n=16, pitch=0.3):
    x = (np.arange(n) - (n - 1) / 2) * pitch
    return np.stack([x, np.zeros(n), np.zeros(n)], 1), np.tile([0.0, 1.0, 0.0], (n, 1))


def phased_2d(n=33, opening_deg=60.0):
    th = np.deg2rad(np.linspace(-opening_deg / 2, opening_deg / 2, n))
    return np.zeros((n, 3)), np.stack([np.sin(th), np.cos(th), np.zeros(n)], 1), th


def convex_2d(n=24, radius=20.0, opening_deg=50.0):
    th = np.deg2rad(np.linspace(-opening_deg / 2, opening_deg / 2, n))
    d = np.stack([np.sin(th), np.cos(th), np.zeros(n)], 1)
    return radius * d - np.array([0.0, radius, 0.0]), d


def pixel_positions(tab):
    sx, sy, sz = tab["size"]
    i = np.arange(sx * sy * sz)
    res, lo = tab["resolution"], tab["bb_min"]
    return np.stack([(i % sx) * res + lo[0], ((i // sx) % sy) * res + lo[1], (i // (sx * sy)) * res + lo[2]], 1)


def test_linear_2d_closed_form():
    pos, dirs = linear_2d()
    depth, ns, res = 30.0, 301, 0.07
    tab = SC.tables(pos, dirs, (16, 1), depth, ns, res)
    x0, x1 = pos[0, 0], pos[-1, 0]
    assert tab["size"] == (int(np.ceil((x1 - x0) / res)) + 1, int(np.ceil(depth / res)) + 1, 1)
    p = pixel_positions(tab)
    inside = tab["index"] != SC.OUTSIDE
    # the scanned region is the rectangle [x0, x1] x [0, depth]
    expect = (p[:, 0] <= x1 + 1e-12) & (p[:, 1] <= depth + 1e-12)
    assert np.array_equal(inside, expect)
    idx = tab["index"][inside].astype(np.int64)
    line, row = idx % 16, idx // 16
    px, py = p[inside, 0], p[inside, 1]
    t = (px - pos[line, 0]) / 0.3
    assert np.all((t >= -1e-12) & (t <= 1 + 1e-12))
    np.testing.assert_allclose(tab["weights"][0][inside], t, atol=2e-7)
    sd = depth / (ns - 1)
    clear = np.abs(py / sd - np.round(py / sd)) > 1e-9
    assert np.array_equal(row[clear], np.floor(py / sd)[clear].astype(np.int64))
    # the reference's 2D depth weight: a length in mm
    np.testing.assert_allclose(tab["weights"][1][inside][clear], (py - row * sd)[clear], atol=1e-6)
    norm = SC.tables(pos, dirs, (16, 1), depth, ns, res, normalized=True)
    np.testing.assert_allclose(norm["weights"][1][inside][clear], (py / sd - row)[clear], atol=1e-6)


def test_phased_2d_closed_form_and_area():
    pos, dirs, th = phased_2d()
    depth, ns = 40.0, 401
    tab = SC.tables(pos, dirs, (33, 1), depth, ns, 0.15)
    p = pixel_positions(tab)
    inside = tab["index"] != SC.OUTSIDE
    idx = tab["index"][inside].astype(np.int64)
    line = idx % 33
    px, py = p[inside, 0], p[inside, 1]
    r, phi = np.hypot(px, py), np.arctan2(px, py)
    dth = th[1] - th[0]
    far = r > 2.0
    # slerp interpolates the angle linearly: t is the angular fraction, d the radius
    assert np.all(phi[far] >= th[line[far]] - 1e-9) and np.all(phi[far] <= th[line[far] + 1] + 1e-9)
    np.testing.assert_allclose(tab["weights"][0][inside][far], ((phi - th[line]) / dth)[far], atol=1e-5 / 2 / dth + 1e-6)
    sd = depth / (ns - 1)
    row = idx // 33
    clear = np.abs(r / sd - np.round(r / sd)) > 1e-6
    assert np.array_equal(row[clear], np.floor(r / sd)[clear].astype(np.int64))
    # scanned area: the fan polygon (apex + the end points), within one pixel row of its perimeter
    area = 0.5 * depth ** 2 * np.sin(dth) * 32
    perim = 2 * depth + 32 * 2 * depth * np.sin(dth / 2)
    assert abs(inside.sum() * 0.15 ** 2 - area) < perim * 0.15


@pytest.mark.parametrize("kind", ["convex", "phased"])
def test_2d_mapping_identity(kind):
    if kind == "convex":
        pos, dirs = convex_2d()
        layout = (24, 1)
    else:
        pos, dirs, _ = phased_2d(9, 40.0)
        layout = (9, 1)
    depth = 25.0
    tab = SC.tables(pos, dirs, layout, depth, 251, 0.2)
    p = pixel_positions(tab)
    inside = np.nonzero(tab["index"] != SC.OUTSIDE)[0]
    line = (tab["index"][inside] % layout[0]).astype(np.int64)
    for i in np.unique(line):
        sel = inside[line == i]
        t, d = SC.map_to_parameters_2d(pos[i], pos[i + 1], dirs[i], dirs[i + 1], p[sel])
        recon = (1 - t)[:, None] * pos[i] + t[:, None] * pos[i + 1] + d[:, None] * SC._slerp(dirs[i], dirs[i + 1], t)
        assert np.abs(recon - p[sel]).max() < 2e-5
        assert np.all((t > -1e-9) & (t < 1 + 1e-9)) and np.all(d <= depth + 1e-6)
        np.testing.assert_allclose(tab["weights"][0][sel], t, atol=1e-7)


def grid_3d(nx=5, ny=4, px=0.4, pz=0.5):
    gx, gz = np.meshgrid((np.arange(nx) - (nx - 1) / 2) * px, (np.arange(ny) - (ny - 1) / 2) * pz)
    pos = np.stack([gx.ravel(), np.zeros(nx * ny), gz.ravel()], 1)
    return pos, np.tile([0.0, 1.0, 0.0], (nx * ny, 1))


def test_linear_3d_closed_form():
    pos, dirs = grid_3d()
    depth, ns, res = 6.0, 61, 0.13
    tab = SC.tables(pos, dirs, (5, 4), depth, ns, res)
    assert tab["is3d"]
    p = pixel_positions(tab)
    inside = tab["index"] != SC.OUTSIDE
    expect = (p[:, 0] <= pos[:, 0].max() + 1e-9) & (p[:, 1] <= depth + 1e-9) & (p[:, 2] <= pos[:, 2].max() + 1e-9)
    assert np.array_equal(inside, expect)
    idx = tab["index"][inside].astype(np.int64)
    line, row = idx % 20, idx // 20
    pi = p[inside]
    np.testing.assert_allclose(tab["weights"][0][inside], (pi[:, 0] - pos[line, 0]) / 0.4, atol=1e-6)
    np.testing.assert_allclose(tab["weights"][2][inside], (pi[:, 2] - pos[line, 2]) / 0.5, atol=1e-6)
    sd = depth / (ns - 1)
    clear = np.abs(pi[:, 1] / sd - np.round(pi[:, 1] / sd)) > 1e-9
    assert np.array_equal(row[clear], np.floor(pi[clear, 1] / sd).astype(np.int64))
    np.testing.assert_allclose(tab["weights"][1][inside][clear], (pi[:, 1] / sd - row)[clear], atol=1e-5)


def convex_3d(radius=15.0):
    ax, el = np.deg2rad(np.linspace(-20, 20, 4)), np.deg2rad(np.linspace(-15, 15, 3))
    A, E = np.meshgrid(ax, el)
    dirs = np.stack([np.sin(A) * np.cos(E), np.cos(A) * np.cos(E), np.sin(E)], -1).reshape(-1, 3)
    return radius * dirs - np.array([0.0, radius, 0.0]), dirs


def test_convex_3d_mapping_identity():
    pos, dirs = convex_3d()
    depth = 8.0
    tab = SC.tables(pos, dirs, (4, 3), depth, 81, 0.25)
    p = pixel_positions(tab)
    inside = np.nonzero(tab["index"] != SC.OUTSIDE)[0]
    assert len(inside) > 1000
    base = (tab["index"][inside] % 12).astype(np.int64)
    for b in np.unique(base):
        q = [b, b + 1, b + 4, b + 5]
        sel = inside[base == b]
        ok, t1, t2, d = SC.map_to_parameters_3d(pos[q], dirs[q], p[sel])
        assert ok.all()
        a, ax_, ay, axy = pos[q]
        bx1 = (1 - t1)[:, None] * a + t1[:, None] * ax_
        bx2 = (1 - t1)[:, None] * ay + t1[:, None] * axy
        line = SC._slerp(SC._slerp(dirs[q[0]], dirs[q[1]], t1), SC._slerp(dirs[q[2]], dirs[q[3]], t1), t2)
        # the pixel lies d along the interpolated scanline from the bilinear base point
        recon = (1 - t2)[:, None] * bx1 + t2[:, None] * bx2 + d[:, None] * line
        assert np.abs(recon - p[sel]).max() < 5e-5
        assert np.all((t1 > -1e-6) & (t1 < 1 + 1e-6) & (t2 > -1e-6) & (t2 < 1 + 1e-6))
        np.testing.assert_allclose(tab["weights"][0][sel], t1, atol=1e-7)
        np.testing.assert_allclose(tab["weights"][2][sel], t2, atol=1e-7)


def test_common_origin_3d_keeps_the_first_estimate():
    """With every scanline starting at one point (a phased 3D scan) the
    iteration's plane normals cross(direction, 0) are undefined, the distance
    is NaN and the loop stops after its first step (ScanConverter.cu:228-236):
    t is the first distance-weighted estimate.  The restatement keeps that."""
    _, dirs = convex_3d()
    pos = np.zeros_like(dirs)
    q = [0, 1, 4, 5]
    x = 6.0 * (dirs[q].sum(axis=0) / np.linalg.norm(dirs[q].sum(axis=0)))[None, :]
    with np.errstate(invalid="ignore", divide="ignore"):
        ok, t1, t2, d = SC.map_to_parameters_3d(pos[q], dirs[q], x)
    assert ok.all() and np.isfinite(t1).all() and np.isfinite(t2).all()
    assert d[0] == pytest.approx(6.0)
    # by symmetry the first estimate is the centre of the quad
    assert t1[0] == pytest.approx(0.5, abs=1e-3) and t2[0] == pytest.approx(0.5, abs=0.05)


def test_convert_reproduces_linear_data():
    pos, dirs = linear_2d(12, 0.5)
    depth, ns = 20.0, 81
    sd = depth / (ns - 1)
    tab = SC.tables(pos, dirs, (12, 1), depth, ns, 0.11, normalized=True)
    # data linear in (x, depth): bilinear interpolation reproduces it
    xs = pos[:, 0]
    frame = (3.0 + 2.0 * xs[None, :] + 0.5 * (np.arange(ns) * sd)[:, None]).astype(np.float32)
    out = SC.convert(tab, frame[None], np.float32)[0]
    p = pixel_positions(tab)
    valid = (tab["index"] != SC.OUTSIDE) & (tab["index"].astype(np.int64) // 12 < ns - 1)
    np.testing.assert_allclose(out[valid], (3.0 + 2.0 * p[valid, 0] + 0.5 * p[valid, 1]), rtol=2e-6, atol=2e-5)
    assert np.all(out[~valid] == 0)
    # the reference's mm depth weight: the depth coordinate becomes (row + wy_mm) * sd
    ref = SC.tables(pos, dirs, (12, 1), depth, ns, 0.11)
    out_ref = SC.convert(ref, frame[None], np.float32)[0]
    row = ref["index"][valid].astype(np.int64) // 12
    want = 3.0 + 2.0 * p[valid, 0] + 0.5 * (row + ref["weights"][1][valid].astype(np.float64)) * sd
    np.testing.assert_allclose(out_ref[valid], want, rtol=2e-6, atol=2e-5)


def test_convert_3d_trilinear_and_types():
    pos, dirs = grid_3d(4, 3, 0.5, 0.5)
    depth, ns = 5.0, 51
    sd = depth / (ns - 1)
    tab = SC.tables(pos, dirs, (4, 3), depth, ns, 0.17)
    frame = (1.0 + pos[:, 0][None, :] - 2.0 * pos[:, 2][None, :] + 0.25 * (np.arange(ns) * sd)[:, None])
    frame = frame.astype(np.float32)
    out = SC.convert(tab, np.stack([frame, 2 * frame]), np.float32)
    p = pixel_positions(tab)
    idx = tab["index"].astype(np.int64)
    valid = (tab["index"] != SC.OUTSIDE) & (idx + 1 + 4 + 12 < ns * 12)
    want = 1.0 + p[valid, 0] - 2.0 * p[valid, 2] + 0.25 * p[valid, 1]
    np.testing.assert_allclose(out[0, valid], want, rtol=2e-6, atol=3e-5)
    np.testing.assert_allclose(out[1, valid], 2 * want, rtol=2e-6, atol=6e-5)
    assert np.all(out[:, ~valid] == 0)
    u8 = SC.convert(tab, (frame * 100)[None], np.uint8)[0]
    f = SC.convert(tab, (frame * 100)[None], np.float32)[0]
    assert np.array_equal(u8, np.clip(f, 0, 255).astype(np.uint8))
    i16 = SC.convert(tab, (frame * -20000)[None], np.int16)[0]
    assert i16.min() == -32767


def test_oracle_flags_a_planted_transposed_weight():
    """A plausible slip, (1 - t) for t, must fail the linear closed form."""
    pos, dirs = linear_2d(6, 0.4)
    tab = SC.tables(pos, dirs, (6, 1), 10.0, 101, 0.09)
    p = pixel_positions(tab)
    inside = tab["index"] != SC.OUTSIDE
    line = tab["index"][inside].astype(np.int64) % 6
    t = (p[inside, 0] - pos[line, 0]) / 0.4
    wrong = 1 - tab["weights"][0][inside]
    assert np.abs(wrong - t).max() > 0.5
    assert np.abs(tab["weights"][0][inside] - t).max() < 1e-6
