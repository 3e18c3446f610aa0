"""The batched stream (supra_batch_*) on the GPU: every acquisition of a
mixed, uneven batch equals the single-acquisition entry points bit for bit
(delay-and-sum: supra_das_run; minimum variance: supra_mv_das_run), on
repeated graph replays with other GPU work in between, and the device
timestamps give a completion time for every acquisition."""
import numpy as np
import pytest

from supra_amd import _native as N
from supra_amd import workloads as W
from supra_amd.beamformer import RxBeamformer, USRawData, mv_das_perform_rx_beamforming

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


SPECS = [  # (num_rx, num_depths, aperture, fs, mode, rf dtype, extra)
    (64, 700, 32, 40e6, "das", "i16", {}),
    (48, 520, 64, 20e6, "mv", "i16", {"temporal_smoothing": 5, "loading_factor": 0.01}),
    (40, 450, 128, 40e6, "mv", "f32", {"temporal_smoothing": 3, "loading_factor": 0.05}),
    (72, 300, 64, 20e6, "das", "f32", {"additional_offset": 3, "window_type": N.WINDOW_HAMMING}),
    (32, 611, 32, 40e6, "mv", "i16", {"temporal_smoothing": 2, "compute_means": True, "loading_factor": 0.1}),
    (56, 333, 64, 40e6, "das", "i16", {"out_dtype": torch.int16}),
]


def build(specs):
    from supra_amd.batch import DAS, MV
    items, refs = [], []
    for i, (n_rx, n_d, ap, fs, mode, dt, extra) in enumerate(specs):
        w = W.linear_acquisition(n_rx, n_d, ap, fs, 60, scatter_seed=50 + i, noise_seed=80 + i)
        g = w.geometry()
        rf = torch.from_numpy(w.rf_host(dtype=np.int16 if dt == "i16" else np.float32)).cuda()
        eng = RxBeamformer(g)
        out_dtype = extra.get("out_dtype", torch.float32)
        out = torch.full((g.num_depths, g.num_rx), -7, dtype=out_dtype, device="cuda")
        it = dict(engine=eng, rf=rf, out=out, fs_hz=fs, mode=DAS if mode == "das" else MV,
                  window_type=extra.get("window_type", N.WINDOW_HANN),
                  additional_offset=extra.get("additional_offset", 0),
                  temporal_smoothing=extra.get("temporal_smoothing", 3),
                  compute_means=extra.get("compute_means", False),
                  loading_factor=extra.get("loading_factor", 1.0))
        raw = USRawData(rf, fs, g)
        if mode == "das":
            ref = eng.perform_rx_beamforming(N.BF_DELAY_AND_SUM, raw, 1.0, 1540e3, it["window_type"], 0.0, False,
                                             it["additional_offset"], out_dtype=out_dtype)
        else:
            ref = mv_das_perform_rx_beamforming(eng, raw, 1.0, 1540e3, it["window_type"], 0.0, 0,
                                                it["temporal_smoothing"], 1.5, it["compute_means"],
                                                it["loading_factor"])
        items.append(it)
        refs.append(ref.data)
    torch.cuda.synchronize()
    return items, [r.cpu().numpy() for r in refs]


@pytest.mark.parametrize("autotune", [False, True])
def test_batch_matches_single_calls_on_replays(device, autotune):
    from supra_amd.batch import Batch
    items, refs = build(SPECS)
    b = Batch(items, autotune=autotune)
    print(b.config())
    for rep in range(3):
        for it in items:                       # eager work between replays
            it["out"].fill_(-7)
        torch.cuda.synchronize()
        b.launch()
        lat, total = b.latencies()
        for k, (it, ref) in enumerate(zip(items, refs)):
            got = it["out"].cpu().numpy()
            same = (got == ref) | (np.isnan(got) & np.isnan(ref))
            assert same.all(), f"replay {rep} acquisition {k}: {(~same).sum()} samples differ"
        assert np.all(lat > 0) and abs(total - lat.max()) < 1e-9
        print(f"replay {rep}: per-acquisition completion (us) {np.round(lat, 1).tolist()}")
    b.close()


def test_batch_rejects_int16_mv_output(device):
    from supra_amd.batch import Batch, MV
    w = W.linear_acquisition(32, 300, 32, 40e6, 10, 1, 2)
    g = w.geometry()
    item = dict(engine=RxBeamformer(g), rf=torch.from_numpy(w.rf_host()).cuda(),
                out=torch.zeros((g.num_depths, g.num_rx), dtype=torch.int16, device="cuda"), fs_hz=40e6, mode=MV)
    with pytest.raises(N.SupraError):
        Batch([item], autotune=False)


def test_batch_rejects_wrong_output_shape(device):
    """The kernels write through the bare output pointer, so the Python
    binding checks the image shape against the engine's geometry first."""
    from supra_amd.batch import Batch, DAS
    w = W.linear_acquisition(32, 300, 32, 40e6, 10, 1, 2)
    g = w.geometry()
    item = dict(engine=RxBeamformer(g), rf=torch.from_numpy(w.rf_host()).cuda(),
                out=torch.zeros((g.num_depths - 1, g.num_rx), dtype=torch.float32, device="cuda"), fs_hz=40e6,
                mode=DAS)
    with pytest.raises(ValueError):
        Batch([item], autotune=False)
