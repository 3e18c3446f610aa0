"""SUPRA's raw-data mock files through the C ABI (host only): .json metadata
(writeMetaDataForMock / readMetaDataForMockJson) and the ASCII .mock v1
layout (readMetaDataForMockAscii) round-trip the receive geometry exactly,
the geometry read back beamforms to the same image, raw frames are counted
and read like UltrasoundInterfaceRawDataMock does, and a hand-written v1 file
with known values parses to them."""
import numpy as np
import pytest

from oracle import oracle as O
from supra_amd import _native as N
from supra_amd import workloads as W
from supra_amd.mock import MockMeta, read_mock, write_frames, write_mock


def meta_of(w, g):
    return MockMeta(g.num_elements, tuple(g.element_layout), w.num_channels, w.num_samples, w.num_tx,
                    tuple(g.rx_layout), g.num_depths, g.depth, w.fs_hz, g.speed_of_sound_mm_per_s)


@pytest.mark.parametrize("suffix", [".json", ".mock"])
@pytest.mark.parametrize("kind", ["linear", "matrix"])
def test_metadata_round_trip(tmp_path, suffix, kind):
    w = W.linear_2d(num_scatterers=10) if kind == "linear" else W.matrix_3d_small(n_el=6, n_lines=5, n_depths=90)
    g = w.geometry()
    path = tmp_path / ("meta" + suffix)
    write_mock(str(path), g, meta_of(w, g))
    g2, m2 = read_mock(str(path))
    assert m2 == meta_of(w, g)
    assert g2.rx_layout == g.rx_layout and g2.element_layout == g.element_layout
    assert g2.scanlines.tobytes() == g.scanlines.tobytes()
    assert np.array_equal(g2.element_x, g.element_x) and np.array_equal(g2.element_y, g.element_y)
    if suffix == ".json":
        # the json reader deduces the depths as depth * i / (N - 1)
        expect = (g.depth * np.arange(g.num_depths, dtype=np.float64) / (g.num_depths - 1)).astype(np.float32)
        assert np.array_equal(g2.depths, expect)
    else:
        assert np.array_equal(g2.depths, g.depths)
    rf = w.rf_host()
    kw = w.das_kwargs()
    g2.depths = g.depths.copy()
    assert np.array_equal(O.das(g2, rf, w.fs_hz, **kw), O.das(g, rf, w.fs_hz, **kw))


def test_raw_frames(tmp_path):
    w = W.linear_acquisition(16, 300, 8, 40e6, 5, 1, 2)
    g = w.geometry()
    meta = meta_of(w, g)
    frames = np.stack([W.linear_acquisition(16, 300, 8, 40e6, 5, 10 + f, 20 + f).rf_host() for f in range(3)])
    raw = tmp_path / "seq.raw"
    write_frames(str(raw), frames)
    import ctypes as C
    m = meta.to_c()
    n = C.c_uint64()
    N.check(N.lib().supra_mock_count_frames(str(raw).encode(), C.byref(m), C.byref(n)))
    assert n.value == 3
    out = np.zeros((2,) + frames.shape[1:], dtype=np.int16)
    N.check(N.lib().supra_mock_read_frames(str(raw).encode(), C.byref(m), 1, 2, out.ctypes.data))
    assert np.array_equal(out, frames[1:])
    assert N.lib().supra_mock_read_frames(str(raw).encode(), C.byref(m), 2, 2, out.ctypes.data) == \
        N.SUPRA_ERR_INVALID_ARG


def test_hand_written_ascii_v1(tmp_path):
    """Two scanlines of a 4-element linear array in the v1 layout (values as
    the reference's writer prints them, 9 significant digits)."""
    # position xyz, direction xyz, maxElementDistance xy, then per transmit
    # first.x first.y last.x last.y txScanlineIdx initialDelay txWeight
    line0 = "-0.15 0 0 0 1 0 0.45 0 " + "0 0 3 0 0 2.5e-07 1 " + "0 0 0 0 0 0 0 " * 3
    line1 = "0.15 0 0 0 1 0 0.45 0 " + "1 0 4 0 1 3.5e-07 1 " + "0 0 0 0 0 0 0 " * 3
    text = ("rawDataMockMetadata v 1\n"
            "4 4 1 4 50 2 2 1 10 40000000 5 1540000\n" + line0 + "\n" + line1 + "\n"
            "0 2.5 5 7.5 10\n"
            "-0.45 -0.15 0.15 0.45\n"
            "0 0 0 0\n")
    p = tmp_path / "tiny.mock"
    p.write_text(text)
    g, m = read_mock(str(p))
    assert (m.num_elements, m.element_layout, m.num_channels, m.num_samples, m.num_tx) == (4, (4, 1), 4, 50, 2)
    assert m.rx_scanline_layout == (2, 1) and m.num_depths == 5
    assert (m.depth, m.sampling_frequency_hz, m.speed_of_sound_mm_per_s) == (10.0, 40e6, 1540e3)
    s = g.scanlines
    assert s["position"][1, 0] == 0.15 and s["max_element_distance"][0, 0] == 0.45
    assert s["tx"][1, 0]["first"].tolist() == [1, 0] and s["tx"][1, 0]["last"].tolist() == [4, 0]
    assert s["tx"][1, 0]["tx_idx"] == 1 and s["tx"][1, 0]["initial_delay"] == 3.5e-7
    assert s["tx_weights"][:, 0].tolist() == [1.0, 1.0] and np.all(s["tx_weights"][:, 1:] == 0)
    assert g.depths.tolist() == [0, 2.5, 5, 7.5, 10]
    assert np.array_equal(g.element_x, np.float32([-0.45, -0.15, 0.15, 0.45]))


def test_errors(tmp_path):
    bad = tmp_path / "meta.txt"
    bad.write_text("{}")
    with pytest.raises(N.SupraError):
        read_mock(str(bad))
    trunc = tmp_path / "t.mock"
    trunc.write_text("rawDataMockMetadata v 1\n4 4 1 4 50 2 2 1 10 40000000 5 1540000\n-0.15 0 0\n")
    with pytest.raises(N.SupraError):
        read_mock(str(trunc))
    js = tmp_path / "b.json"
    js.write_text("{ \"numElements\" : 4, ")
    with pytest.raises(N.SupraError):
        read_mock(str(js))
