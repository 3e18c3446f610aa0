"""Scan conversion on the GPU through the C ABI against the numpy oracle:
the library's tables (2D built on the host, 3D built by GPU kernels) equal
the oracle's (mask, sample index; weights within 1e-6, the float rounding of
double-precision parameters), the conversion kernel is the oracle's float
arithmetic bit for bit on the library's own tables for every input/output
type, batches of frames equal single frames, the node rebuilds its tables
only when the scan changes, and the chain DAS -> scan conversion runs on a
beamformed linear-array frame."""
import types

import numpy as np
import pytest

from oracle import scanconv as SC
from supra_amd import _native as N
from supra_amd.geometry import SCANLINE_DTYPE, BeamformerConfig, RxGeometry

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def scan(positions, directions, layout, depth, num_samples):
    s = np.zeros(len(positions), SCANLINE_DTYPE)
    s["position"] = positions
    s["direction"] = directions
    return types.SimpleNamespace(scanlines=s, rx_layout=tuple(layout), depth=float(depth), num_depths=num_samples)


def geometries():
    n = 24
    x = (np.arange(n) - (n - 1) / 2) * 0.3
    lin = scan(np.stack([x, 0 * x, 0 * x], 1), np.tile([0.0, 1.0, 0.0], (n, 1)), (n, 1), 20.0, 401)
    th = np.deg2rad(np.linspace(-35, 35, 41))
    d = np.stack([np.sin(th), np.cos(th), 0 * th], 1)
    phased = scan(np.zeros((41, 3)), d, (41, 1), 30.0, 301)
    th = np.deg2rad(np.linspace(-25, 25, 32))
    d = np.stack([np.sin(th), np.cos(th), 0 * th], 1)
    convex = scan(20.0 * d - np.array([0.0, 20.0, 0.0]), d, (32, 1), 25.0, 251)
    ax, el = np.deg2rad(np.linspace(-20, 20, 6)), np.deg2rad(np.linspace(-15, 15, 5))
    A, E = np.meshgrid(ax, el)
    d3 = np.stack([np.sin(A) * np.cos(E), np.cos(A) * np.cos(E), np.sin(E)], -1).reshape(-1, 3)
    convex3d = scan(15.0 * d3 - np.array([0.0, 15.0, 0.0]), d3, (6, 5), 10.0, 101)
    gx, gz = np.meshgrid((np.arange(5) - 2) * 0.4, (np.arange(4) - 1.5) * 0.5)
    lin3d = scan(np.stack([gx.ravel(), 0 * gx.ravel(), gz.ravel()], 1), np.tile([0.0, 1.0, 0.0], (20, 1)), (5, 4),
                 6.0, 61)
    return {"linear": (lin, 0.05), "phased": (phased, 0.12), "convex": (convex, 0.1),
            "convex3d": (convex3d, 0.2), "linear3d": (lin3d, 0.13)}


GEOMS = geometries()


def oracle_tables(g, res, normalized=False):
    return SC.tables(g.scanlines["position"], g.scanlines["direction"], g.rx_layout, g.depth, g.num_depths, res,
                     normalized=normalized)


def compare_tables(sc, tab, max_mask_diff=0):
    mask, idx, w = sc.tables()
    assert sc.image_size == tab["size"]
    np.testing.assert_allclose(sc.bb_min, tab["bb_min"], rtol=0, atol=1e-12)
    o_in = tab["index"] != SC.OUTSIDE
    differ = (mask.astype(bool) != o_in).sum()
    assert differ <= max_mask_diff, f"{differ} mask pixels differ"
    both = mask.astype(bool) & o_in
    same_idx = idx[both] == tab["index"][both]
    # a sample index may differ only where the depth weight sits on a sample boundary
    wy = np.minimum(w[1][both], tab["weights"][1][both])
    assert np.all(same_idx | (np.abs(wy) < 1e-5) | (np.abs(1 - np.maximum(w[1][both], tab["weights"][1][both])) < 1e-5))
    k = same_idx
    for a in (0, 1, 2):
        np.testing.assert_allclose(w[a][both][k], tab["weights"][a][both][k], rtol=0, atol=2e-6)
    return mask, idx, w, both.sum()


@pytest.mark.parametrize("name", list(GEOMS))
def test_tables_match_oracle(device, name):
    from supra_amd.scanconv import ScanConverter
    g, res = GEOMS[name]
    sc = ScanConverter()
    sc.update_internals(g, resolution=res)
    tab = oracle_tables(g, res)
    # 3D membership uses det expansions in two orders: allow a handful of face voxels
    _, _, _, n = compare_tables(sc, tab, max_mask_diff=0 if "3d" not in name else max(3, tab["index"].size // 20000))
    assert n > 500
    print(f"[scanconv] {name}: {sc.image_size} pixels, {n} inside, tables match the oracle")
    m = sc.get_mask()
    torch.cuda.synchronize()
    assert np.array_equal(m.cpu().numpy().ravel(), sc.tables()[0])
    sc.close()


def test_normalized_depth_weight_flag(device):
    from supra_amd.scanconv import ScanConverter
    g, res = GEOMS["phased"]
    sc = ScanConverter()
    sc.update_internals(g, resolution=res, normalized_depth_weight=True)
    compare_tables(sc, oracle_tables(g, res, normalized=True))


@pytest.mark.parametrize("name,in_dt,out_dt", [("linear", "f32", "f32"), ("phased", "i16", "u8"),
                                                ("convex", "u8", "i16"), ("convex3d", "f32", "i16"),
                                                ("linear3d", "i16", "f32"), ("phased", "f32", "u8")])
def test_convert_bit_exact_on_library_tables(device, name, in_dt, out_dt):
    """The kernel is the oracle's float arithmetic on the same tables, for a
    batch of three frames; each frame also equals a single-frame call."""
    from supra_amd.scanconv import ScanConverter
    g, res = GEOMS[name]
    sc = ScanConverter()
    sc.update_internals(g, resolution=res)
    mask, idx, w = sc.tables()
    tab = dict(size=sc.image_size, index=idx, weights=w, layout=g.rx_layout, is3d=sc.is_3d,
               num_samples=g.num_depths)
    n_rx = g.rx_layout[0] * g.rx_layout[1]
    rng = np.random.default_rng(len(name) * 7 + len(in_dt))
    npt = {"f32": np.float32, "i16": np.int16, "u8": np.uint8}
    tt = {"f32": torch.float32, "i16": torch.int16, "u8": torch.uint8}
    if in_dt == "u8":
        frames = rng.integers(0, 256, (3, g.num_depths, n_rx)).astype(np.uint8)
    else:
        frames = (rng.normal(size=(3, g.num_depths, n_rx)) * (9000 if out_dt != "f32" else 1)).astype(npt[in_dt])
    if in_dt == "f32":
        frames[0, 5, 3] = np.nan
        frames[1, 7, 1] = np.inf
    dev = torch.from_numpy(frames).cuda()
    got = sc.convert(dev, tt[out_dt])
    single = sc.convert(dev[1], tt[out_dt])
    torch.cuda.synchronize()
    want = SC.convert(tab, frames, npt[out_dt])
    got = got.cpu().numpy().reshape(3, -1)
    assert got.dtype == want.dtype
    same = (got == want) | (np.isnan(got) & np.isnan(want))
    assert same.all(), f"{(~same).sum()} of {got.size} pixels differ"
    assert np.array_equal(single.cpu().numpy().ravel(), got[1], equal_nan=True)
    # an output one element into a larger buffer is not aligned to the packed
    # stores of 4 pixels: the per-pixel kernel writes it, with the same values
    shape = (3,) + tuple(single.shape)
    buf = torch.zeros(int(np.prod(shape)) + 1, dtype=tt[out_dt], device="cuda")
    shifted = sc.convert(dev, tt[out_dt], out=buf[1:].view(shape))
    torch.cuda.synchronize()
    assert np.array_equal(shifted.cpu().numpy().reshape(3, -1), got, equal_nan=True)
    # outside the scan the image is 0
    assert np.all(got[:, mask == 0] == 0)


def test_node_and_beamformer_chain(device):
    """BeamformingNode's image -> ScanConverterNode: tables built once per
    scan, resolution forced or from the sample distance, float image equal
    to the oracle's conversion of the beamformed frame."""
    from supra_amd.beamformer import RxBeamformer, USRawData
    from supra_amd.scanconv import ScanConverterNode
    from supra_amd import workloads as W
    w = W.linear_2d(num_scatterers=40)
    g = w.geometry()
    rf = torch.from_numpy(w.rf_host()).cuda()
    bf = RxBeamformer(g)
    kw = w.das_kwargs()
    img = bf.perform_rx_beamforming(N.BF_DELAY_AND_SUM, USRawData(rf, w.fs_hz, g), kw["f_number"],
                                    kw["speed_of_sound_mm_per_s"], kw["window_type"], 0.0, False, 0)
    node = ScanConverterNode(imageResolutionForced=True, imageResolution=0.1)
    out = node.process(img)
    first = node.converter._sc.value
    out2 = node.process(img)
    assert node.converter._sc.value == first          # same scan: tables kept
    torch.cuda.synchronize()
    assert out.properties["imageState"] == "Scan"
    sx, sy, _ = out.size
    assert tuple(out.data.shape) == (sy, sx)
    tab = SC.tables(g.scanlines["position"], g.scanlines["direction"], g.rx_layout, g.depth, g.num_depths, 0.1)
    assert (sx, sy, 1) == tab["size"]
    mask, idx, wts = node.converter.tables()
    mine = dict(size=tab["size"], index=idx, weights=wts, layout=g.rx_layout, is3d=False, num_samples=g.num_depths)
    frame = img.data.cpu().numpy()
    want = SC.convert(mine, frame[None], np.float32)[0]
    got = out.data.cpu().numpy().ravel()
    assert np.array_equal(got, want, equal_nan=True)
    assert np.array_equal(out2.data.cpu().numpy().ravel(), got, equal_nan=True)
    assert tuple(node.mask.shape) == (sy, sx)
    node.set("imageResolution", 0.2)
    out3 = node.process(img)
    assert node.converter._sc.value != first and out3.size[0] < sx
    bf.close()


def test_scanconv_rejects_bad_arguments(device):
    from supra_amd.scanconv import ScanConverter
    g, res = GEOMS["linear"]
    sc = ScanConverter()
    sc.update_internals(g, resolution=res)
    bad = torch.zeros((g.num_depths + 1, 24), device="cuda")
    with pytest.raises(ValueError):
        sc.convert(bad)
    lib = N.lib()
    out = torch.zeros(sc.num_pixels, device="cuda")
    ok = torch.zeros((g.num_depths, 24), device="cuda", dtype=torch.float64)
    assert lib.supra_scanconv_run(sc._sc, ok.data_ptr(), 9, 1, out.data_ptr(), N.SUPRA_FLOAT32, None) == \
        N.SUPRA_ERR_UNSUPPORTED
    assert lib.supra_scanconv_run(sc._sc, ok.data_ptr(), N.SUPRA_FLOAT32, 0, out.data_ptr(), N.SUPRA_FLOAT32,
                                  None) == N.SUPRA_OK
