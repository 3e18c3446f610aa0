"""DelayAndStdDev and CoherenceFactorDelayAndSum (beamformerType 1 and 2) on
the GPU through the C ABI, bit for bit against the FP32 oracle: the 2D
linear-array frame, int16 images, a sparse element map, transmit
interpolation on a subdivided scan, and a 3D matrix probe; and the node's
dispatch of the XML beamformerType names."""
from dataclasses import replace

import numpy as np
import pytest

from oracle import oracle as O
from supra_amd import _native as N
from supra_amd import workloads as W
from supra_amd.beamformer import BeamformingNode, RxBeamformer, USRawData

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

KINDS = [N.BF_DELAY_AND_STD_DEV, N.BF_COHERENCE_FACTOR]


def run(g, rf, fs, kind, out_dtype=None, **kw):
    bf = RxBeamformer(g)
    img = bf.perform_rx_beamforming(kind, USRawData(torch.from_numpy(np.ascontiguousarray(rf)).cuda(), fs, g),
                                    kw["f_number"], kw["speed_of_sound_mm_per_s"], kw["window_type"],
                                    kw.get("window_parameter", 0.0), kw.get("interpolate_transmits", False),
                                    kw.get("additional_offset", 0), out_dtype=out_dtype)
    torch.cuda.synchronize()
    out = img.data.cpu().numpy()
    bf.close()
    return out


def same(a, b, label):
    ok = (a == b) | (np.isnan(a) & np.isnan(b))
    assert ok.all(), f"{label}: {(~ok).sum()} of {a.size} pixels differ"


@pytest.mark.parametrize("kind", KINDS)
def test_variants_linear_2d_bit_exact(device, kind):
    w = W.linear_2d(num_scatterers=60)
    g = w.geometry()
    rf = w.rf_host()
    kw = w.das_kwargs()
    same(run(g, rf, w.fs_hz, kind, **kw), O.das(g, rf, w.fs_hz, sample_beamformer=kind, **kw), f"kind {kind}")
    kw2 = dict(kw, additional_offset=-4, window_type=N.WINDOW_HAMMING)
    got = run(g, rf, w.fs_hz, kind, out_dtype=torch.int16, **kw2)
    ref = O.das(g, rf, w.fs_hz, sample_beamformer=kind, out_int16=True, **kw2).astype(np.int16)
    same(got, ref, f"kind {kind} int16")


@pytest.mark.parametrize("kind", KINDS)
def test_variants_sparse_map_and_interp_tx(device, kind):
    w = W.linear_2d(num_scatterers=30)
    g = w.geometry()
    emap = (np.arange(128) % 64).astype(np.int32)
    emap[np.random.default_rng(4).choice(128, 11, replace=False)] = -1
    g.element_to_channel = emap
    rf = w.rf_host(dtype=np.float32)
    kw = w.das_kwargs()
    same(run(g, rf, w.fs_hz, kind, **kw), O.das(g, rf, w.fs_hz, sample_beamformer=kind, **kw), "sparse map")
    wi = W.linear_2d(num_scatterers=30, interpolate_transmits=True)
    wi = replace(wi, bf=replace(wi.bf, num_scanlines=(64, 1), rx_scanline_subdivision=(2, 1)))
    gi = wi.geometry()
    rfi = wi.rf_host()
    kwi = wi.das_kwargs()
    same(run(gi, rfi, wi.fs_hz, kind, **kwi), O.das(gi, rfi, wi.fs_hz, sample_beamformer=kind, **kwi), "interp tx")


@pytest.mark.parametrize("kind", KINDS)
def test_variants_3d_bit_exact(device, kind):
    w = W.matrix_3d_small(n_el=8, n_lines=6, n_depths=140, n_samples=800)
    g = w.geometry()
    rf = w.rf_host()
    kw = w.das_kwargs()
    same(run(g, rf, w.fs_hz, kind, **kw), O.das(g, rf, w.fs_hz, sample_beamformer=kind, **kw), f"3D kind {kind}")


def test_node_dispatches_variant_names(device):
    w = W.linear_2d(num_scatterers=20)
    g = w.geometry()
    rf = w.rf_host()
    for name, kind in [("DelayAndStdDev", N.BF_DELAY_AND_STD_DEV),
                       ("CoherenceFactorDelayAndSum", N.BF_COHERENCE_FACTOR)]:
        node = BeamformingNode(beamformerType=name, windowType="Hann")
        img = node.process(USRawData(torch.from_numpy(rf).cuda(), w.fs_hz, g))
        torch.cuda.synchronize()
        same(img.data.cpu().numpy(), O.das(g, rf, w.fs_hz, sample_beamformer=kind, **w.das_kwargs()), name)
