"""Scan-conversion oracle: a numpy restatement of SUPRA's ScanConverter
(src/SupraLib/Beamformer/ScanConverter.cu), pixel-vectorised.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the supra_amd package.

* ``image_geometry``  — bounding box and image size (ScanConverter.cu:612-633,
  USImageProperties.cpp:326-340);
* ``tables``          — per-pixel sample index and interpolation weights:
  2D pairs of scanlines, two triangles each (:674-733) and
  mapToParameters2D (:847-895); 3D quads, five tetrahedra (:77-81, :125-172)
  and mapToParameters3D (:183-265); the later pair/quad wins a shared pixel;
* ``convert``         — scanConvert2D / scanConvert3D (:350-428) in float32,
  one rounding per operation, clampCast to the output type.

Conventions of the product that the reference leaves open, restated here:
the 2D kernel's next-row guard (the 3D kernel's, :413), and the optional
normalised 2D depth weight (``normalized=True``).  Double precision
throughout the geometry.  Pinning: the closed forms of parallel (linear
array) and common-origin (phased array) scanline fans, and the defining
identity pixel = base(t) + d * direction(t) for every mapped pixel.
"""
from __future__ import annotations

import numpy as np

EPS = 2.2204e-16              # M_EPS (utilities/utility.h:43)
MAX_ITERATIONS = 1000         # ScanConverter.h:28
DISTANCE_THRESHOLD = 1e-5     # ScanConverter.h:29
TETRAHEDRON_THRESHOLD = 1e-9  # ScanConverter.cu:46
OUTSIDE = np.uint32(0xFFFFFFFF)


def _dot(a, b):
    return (a * b).sum(axis=-1)


def _norm(a):
    return np.sqrt(_dot(a, a))


def _slerp(a, b, t):
    """slerp3 (src/SupraLib/vec.h:424-445), rows of a, b (or broadcast) and t."""
    a = np.broadcast_to(a, np.broadcast_shapes(np.shape(a), np.shape(t)[:1] + (3,)))
    b = np.broadcast_to(b, a.shape)
    with np.errstate(invalid="ignore", divide="ignore"):
        omega = np.arccos(_dot(a, b))
        s = np.sin(omega)
        mid = (np.sin((1.0 - t) * omega) / s)[:, None] * a + (np.sin(t * omega) / s)[:, None] * b
    out = np.where(((omega < EPS) | (t < EPS))[:, None], a, np.where((1.0 - t < EPS)[:, None], b, mid))
    return out


def image_geometry(positions, directions, depth, num_samples, resolution=0.0):
    """(size[3], bb_min, bb_max, resolution, sample_distance)."""
    p = np.asarray(positions, np.float64).reshape(-1, 3)
    d = np.asarray(directions, np.float64).reshape(-1, 3)
    pts = np.concatenate([p, p + depth * d])
    lo, hi = pts.min(axis=0), pts.max(axis=0)
    sd = depth / (num_samples - 1)
    res = resolution if resolution > 0 else sd
    size = (np.ceil((hi - lo) / res) + 1).astype(np.int64)
    return size, lo, hi, res, sd


def _line_conn(a, da, x):
    c = x - a
    return c - _dot(da, c)[..., None] * da


def map_to_parameters_2d(a, b, da, db, x):
    """mapToParameters2D for many points x [n,3]: (t, d), float64."""
    n = len(x)
    low_t, high_t = np.zeros(n), np.ones(n)
    low_c = _line_conn(a, da, x)
    high_c = _line_conn(b, db, x)
    low_d, high_d = _norm(low_c), _norm(high_c)
    t = np.full(n, 0.5)
    base = np.zeros((n, 3))
    dist = np.full(n, 1e10)
    hz = np.all(high_c == 0, axis=1)
    lz = np.all(low_c == 0, axis=1) & ~hz
    active = ~(hz | lz)
    for _ in range(MAX_ITERATIONS):
        act = active & (dist > DISTANCE_THRESHOLD)
        if not act.any():
            break
        i = np.nonzero(act)[0]
        ti = (1 - high_d[i] / (high_d[i] + low_d[i])) * high_t[i] + (1 - low_d[i] / (high_d[i] + low_d[i])) * low_t[i]
        bi = (1 - ti)[:, None] * a + ti[:, None] * b
        ci = _line_conn(bi, _slerp(da, db, ti), x[i])
        di = _norm(ci)
        go_high = _dot(low_c[i], ci) < 0
        h, l = i[go_high], i[~go_high]
        high_t[h], high_c[h], high_d[h] = ti[go_high], ci[go_high], di[go_high]
        low_t[l], low_c[l], low_d[l] = ti[~go_high], ci[~go_high], di[~go_high]
        t[i], base[i], dist[i] = ti, bi, di
    d = _norm(x - base)
    t = np.where(hz, 1.0, np.where(lz, 0.0, t))
    d = np.where(hz, _norm(x - b), np.where(lz, _norm(x - a), d))
    return t, d


def _in_triangle(a, b, c, p):
    """pointInsideTriangle (ScanConverter.cu:818-834) for points p [n,>=2]."""
    def bary(u, v):   # barycentricCoordinate2D(u, v, p)
        return (v[0] - u[0]) * (p[:, 1] - u[1]) - (v[1] - u[1]) * (p[:, 0] - u[0])
    return (bary(b, c) >= 0) & (bary(c, a) >= 0) & (bary(a, b) >= 0)


def _det4(a, b, c, p):
    """barycentricCoordinate3D (ScanConverter.cu:148-172), rows broadcast."""
    a, b, c, p = (np.asarray(v, np.float64) for v in (a, b, c, p))
    m = np.stack(np.broadcast_arrays(a, b, c, p), axis=-2)          # [..., 4, 3]
    ones = np.ones(m.shape[:-1] + (1,))
    return np.linalg.det(np.concatenate([m, ones], axis=-1))


def _in_tetrahedron(a, b, c, d, p):
    w0 = _det4(a, b, c, d)
    if not w0 > 0:
        return np.zeros(len(p), dtype=bool)
    thr = -TETRAHEDRON_THRESHOLD
    return (_det4(p, b, c, d) >= thr) & (_det4(a, p, c, d) >= thr) & (_det4(a, b, p, d) >= thr) & \
        (_det4(a, b, c, p) >= thr)


def _plane_conn(a, na, x):
    return _dot(na, x - a)[..., None] * na


def _normalize(v):
    return v / _norm(v)[..., None]


def map_to_parameters_3d(P, D, x):
    """mapToParameters3D for points x [n,3]; P, D: positions/directions of
    the quad's scanlines (x,y), (x+1,y), (x,y+1), (x+1,y+1).
    Returns (ok, t1, t2, d)."""
    a, ax, ay, axy = P
    da, dax, day, daxy = D
    n = len(x)
    n_xlow = _normalize(np.cross(da, (ay + day) - a))
    n_ylow = _normalize(np.cross((ax + dax) - a, da))
    n_xhigh = _normalize(np.cross(dax, (axy + daxy) - ax))
    n_yhigh = _normalize(np.cross((axy + daxy) - ay, day))
    lo_t, hi_t = np.zeros((n, 2)), np.ones((n, 2))
    lo_cx, hi_cx = _plane_conn(a, n_xlow, x), _plane_conn(ax, n_xhigh, x)
    lo_cy, hi_cy = _plane_conn(a, n_ylow, x), _plane_conn(ay, n_yhigh, x)
    lo_d = np.stack([_norm(lo_cx), _norm(lo_cy)], axis=1)
    hi_d = np.stack([_norm(hi_cx), _norm(hi_cy)], axis=1)
    ok = ~((_dot(lo_cx, hi_cx) > 0) | (_dot(lo_cy, hi_cy) > 0))
    t = np.full((n, 2), 0.5)
    dist = np.full((n, 2), 1e10)
    bx1, bx2 = np.zeros((n, 3)), np.zeros((n, 3))
    for _ in range(MAX_ITERATIONS):
        act = ok & ((dist[:, 0] > DISTANCE_THRESHOLD) | (dist[:, 1] > DISTANCE_THRESHOLD))
        if not act.any():
            break
        i = np.nonzero(act)[0]
        ti = (1 - hi_d[i] / (hi_d[i] + lo_d[i])) * hi_t[i] + (1 - lo_d[i] / (hi_d[i] + lo_d[i])) * lo_t[i]
        t0, t1 = ti[:, 0:1], ti[:, 1:2]
        b1 = (1 - t0) * a + t0 * ax
        b2 = (1 - t0) * ay + t0 * axy
        by1 = (1 - t1) * a + t1 * ay
        by2 = (1 - t1) * ax + t1 * axy
        line = _slerp(_slerp(da, dax, ti[:, 0]), _slerp(day, daxy, ti[:, 0]), ti[:, 1])
        nx = _normalize(np.cross(line, b2 - b1))
        ny = _normalize(np.cross(by2 - by1, line))
        cx = _plane_conn(b1, nx, x[i])
        cy = _plane_conn(by1, ny, x[i])
        dx, dy = _norm(cx), _norm(cy)
        for k, (c, dk, lo_c, hi_c) in enumerate(((cx, dx, lo_cx, hi_cx), (cy, dy, lo_cy, hi_cy))):
            up = _dot(hi_c[i], c) > EPS
            down = ~up & (_dot(lo_c[i], c) > EPS)
            iu, idn = i[up], i[down]
            hi_t[iu, k], hi_c[iu], hi_d[iu, k] = ti[up, k], c[up], dk[up]
            lo_t[idn, k], lo_c[idn], lo_d[idn, k] = ti[down, k], c[down], dk[down]
        t[i], dist[i, 0], dist[i, 1], bx1[i], bx2[i] = ti, dx, dy, b1, b2
    base = (1 - t[:, 1:2]) * bx1 + t[:, 1:2] * bx2
    return ok, t[:, 0], t[:, 1], _norm(x - base)


def tables(positions, directions, layout, depth, num_samples, resolution=0.0, normalized=False):
    """dict(size, bb_min, resolution, sample_distance, is3d, index [pixels]
    uint32 (OUTSIDE off the scan), weights [3][pixels] float32)."""
    lx, ly = int(layout[0]), int(layout[1])
    P = np.asarray(positions, np.float64).reshape(ly, lx, 3)
    D = np.asarray(directions, np.float64).reshape(ly, lx, 3)
    size, lo, hi, res, sd = image_geometry(P, D, depth, num_samples, resolution)
    is3d = not (lx == 1 or ly == 1)
    n_pix = int(np.prod(size))
    n_rx = lx * ly
    index = np.full(n_pix, OUTSIDE, dtype=np.uint32)
    w = np.zeros((3, n_pix), dtype=np.float32)

    def box(points, k):
        a0 = max(0, int(np.floor((points[:, k].min() - lo[k]) / res)))
        a1 = min(int(size[k]) - 1, int(np.ceil((points[:, k].max() - lo[k]) / res)))
        return a0, a1

    if not is3d:
        assert lx >= 2, "a 2D scan runs along x"
        for y in range(ly):
            for x in range(lx - 1):
                s0, s1 = P[y, x], P[y, x + 1]
                e0, e1 = P[y, x] + depth * D[y, x], P[y, x + 1] + depth * D[y, x + 1]
                corners = np.stack([s0, s1, e0, e1])
                (x0, x1), (y0, y1) = box(corners, 0), box(corners, 1)
                if x1 < x0 or y1 < y0:
                    continue
                px, py = np.meshgrid(np.arange(x0, x1 + 1), np.arange(y0, y1 + 1), indexing="ij")
                px, py = px.ravel(), py.ravel()
                pts = np.stack([px * res + lo[0], py * res + lo[1], np.zeros(len(px))], axis=1)
                inside = _in_triangle(e1, e0, s0, pts) | _in_triangle(s0, s1, e1, pts)
                if not inside.any():
                    continue
                px, py, pts = px[inside], py[inside], pts[inside]
                t, d = map_to_parameters_2d(P[y, x], P[y, x + 1], D[y, x], D[y, x + 1], pts)
                fl = np.floor(d / sd)
                wy = d / sd - fl if normalized else d - fl * sd
                o = px + py * int(size[0])
                index[o] = (fl.astype(np.uint64) * n_rx + x + y * lx).astype(np.uint32)
                w[0, o] = t.astype(np.float32)
                w[1, o] = wy.astype(np.float32)
    else:
        for y in range(ly - 1):
            for x in range(lx - 1):
                Pq = np.stack([P[y, x], P[y, x + 1], P[y + 1, x], P[y + 1, x + 1]])
                Dq = np.stack([D[y, x], D[y, x + 1], D[y + 1, x], D[y + 1, x + 1]])
                s, e = Pq, Pq + depth * Dq
                corners = np.concatenate([s, e])
                rng = [box(corners, k) for k in range(3)]
                if any(r[1] < r[0] for r in rng):
                    continue
                vx, vy, vz = (g.ravel() for g in np.meshgrid(*(np.arange(r[0], r[1] + 1) for r in rng),
                                                               indexing="ij"))
                pts = np.stack([vx * res + lo[0], vy * res + lo[1], vz * res + lo[2]], axis=1)
                inside = (_in_tetrahedron(s[0], s[1], s[2], e[0], pts) | _in_tetrahedron(s[1], s[3], s[2], e[3], pts) |
                          _in_tetrahedron(s[1], e[0], e[1], e[3], pts) | _in_tetrahedron(s[2], e[2], e[0], e[3], pts) |
                          _in_tetrahedron(s[1], s[2], e[0], e[3], pts))
                if not inside.any():
                    continue
                vx, vy, vz, pts = vx[inside], vy[inside], vz[inside], pts[inside]
                ok, t1, t2, d = map_to_parameters_3d(Pq, Dq, pts)
                vx, vy, vz, t1, t2, d = vx[ok], vy[ok], vz[ok], t1[ok], t2[ok], d[ok]
                r = d / sd
                fl = np.floor(r)
                o = vx + vy * int(size[0]) + vz * int(size[0]) * int(size[1])
                index[o] = (fl.astype(np.uint64) * n_rx + x + y * lx).astype(np.uint32)
                w[0, o] = t1.astype(np.float32)
                w[1, o] = (r - fl).astype(np.float32)
                w[2, o] = t2.astype(np.float32)
    return dict(size=tuple(int(v) for v in size), bb_min=lo, bb_max=hi, resolution=res, sample_distance=sd,
                is3d=is3d, index=index, weights=w, layout=(lx, ly), num_samples=int(num_samples))


def clamp_cast(v: np.ndarray, dtype) -> np.ndarray:
    """clampCast (utilities/cudaUtility.h:111-147)."""
    lo, hi = {np.dtype(np.int16): (-32767.0, 32767.0), np.dtype(np.uint8): (0.0, 255.0),
              np.dtype(np.float32): (-3.4028234663852886e38, 3.4028234663852886e38)}[np.dtype(dtype)]
    c = np.fmin(np.fmax(v, np.float32(lo)), np.float32(hi))
    c = np.where(np.isnan(v), np.float32(lo), c)
    return c.astype(dtype)


def convert(tab, frames: np.ndarray, out_dtype=np.float32) -> np.ndarray:
    """scanConvert2D / scanConvert3D for frames [F][num_samples][scanlines]:
    [F][pixels] of out_dtype, float32 arithmetic in the kernel's order."""
    lx, ly = tab["layout"]
    n, nx = lx * ly, lx
    f = np.asarray(frames).reshape(len(frames), -1).astype(np.float32)
    total = f.shape[1]
    idx = tab["index"].astype(np.int64)
    inside = tab["index"] != OUTSIDE
    valid = inside & ((idx + 1 + (nx if tab["is3d"] else 0) + n) < total)
    s = np.where(valid, idx, 0)
    one = np.float32(1)
    wx, wy, wz = (np.where(valid, tab["weights"][k], np.float32(0)) for k in range(3))
    g = lambda off: f[:, s + off]   # noqa: E731
    if not tab["is3d"]:
        val = (one - wy) * ((one - wx) * g(0) + wx * g(1)) + wy * ((one - wx) * g(n) + wx * g(1 + n))
    else:
        val = (one - wy) * ((one - wz) * ((one - wx) * g(0) + wx * g(1)) + wz * ((one - wx) * g(nx) + wx * g(1 + nx))) + \
            wy * ((one - wz) * ((one - wx) * g(n) + wx * g(1 + n)) +
                  wz * ((one - wx) * g(nx + n) + wx * g(1 + nx + n)))
    val = np.where(valid[None, :], val, np.float32(0)).astype(np.float32)
    return clamp_cast(val, out_dtype)
