"""Differentiable 2D DAS on the GPU through the C ABI: the per-frame-c forward
against the FP32 oracle (bit-exact) and the fixed-c entry point, the adjoint
against the FP32 oracle's reverse mode (same arithmetic, only the summation
order of the scatter differs) and the FP64 reverse mode, run-to-run
determinism, the torch autograd wrapper, and a plain PyTorch fp32 form of
the same operation."""
from dataclasses import replace

import numpy as np
import pytest

from oracle import oracle as O
from supra_amd import _native as N
from supra_amd import workloads as W
from supra_amd.beamformer import RxBeamformer, USRawData

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

C_MPS = [1450.0, 1540.0, 1597.5]


def frames_rf(w, n, seed0=10):
    """n frames of float32 RF, each from its own scatterer draw."""
    out = []
    for k in range(n):
        out.append(replace(w, scatter_seed=seed0 + k, noise_seed=seed0 + 100 + k).rf_host(dtype=np.float32))
    return np.stack(out)


def make(w, **kw):
    from supra_amd.differentiable import DifferentiableDAS
    a = w.das_kwargs()
    a.update(kw)
    return DifferentiableDAS(w.geometry(), w.fs_hz, a["f_number"], a["window_type"], a["window_parameter"],
                             a["interpolate_transmits"], a["additional_offset"])


def oracle_args(w, **kw):
    a = w.das_kwargs()
    a.update(kw)
    a.pop("speed_of_sound_mm_per_s")
    return a


@pytest.mark.parametrize("interp", [False, True])
def test_forward_per_frame_c_bit_exact(device, interp):
    w = W.linear_2d(num_scatterers=80, interpolate_transmits=interp)
    if interp:
        w = replace(w, bf=replace(w.bf, num_scanlines=(64, 1), rx_scanline_subdivision=(2, 1)))
    g = w.geometry()
    rf = frames_rf(w, len(C_MPS))
    dd = make(w)
    x = torch.from_numpy(rf).cuda()
    c = torch.tensor(C_MPS, dtype=torch.float64, device="cuda")
    img, norm = dd.forward_raw(x, c)
    torch.cuda.synchronize()
    img, norm = img.cpu().numpy(), norm.cpu().numpy()
    a = oracle_args(w)
    bf = RxBeamformer(g)
    for f, cm in enumerate(C_MPS):
        ref = O.das(g, rf[f], w.fs_hz, speed_of_sound_mm_per_s=cm * 1000.0, **a)
        assert np.array_equal(img[f], ref), f"frame {f}: {(img[f] != ref).sum()} pixels differ"
        one = bf.perform_rx_beamforming(N.BF_DELAY_AND_SUM, USRawData(x[f].contiguous(), w.fs_hz, g), a["f_number"],
                                        cm * 1000.0, a["window_type"], a["window_parameter"], interp,
                                        a["additional_offset"])
        torch.cuda.synchronize()
        assert np.array_equal(one.data.cpu().numpy(), img[f])
        _, _, nref = O.das_adjoint(g, rf[f], w.fs_hz, a["f_number"], cm * 1000.0, a["window_type"],
                                   np.zeros_like(ref, dtype=np.float64), want_rf=False, want_norm=True, precision=32,
                                   interpolate_transmits=interp)
        assert np.array_equal(norm[f], nref[0])
    bf.close()
    dd.close()


def run_adjoint(w, rf, c_mps, gi, **kw):
    dd = make(w, **kw)
    x = torch.from_numpy(rf).cuda()
    c = torch.tensor(c_mps, dtype=torch.float64, device="cuda")
    _, norm = dd.forward_raw(x, c)
    grad_rf, grad_c = dd.adjoint_raw(x, c, norm, torch.from_numpy(gi.astype(np.float32)).cuda())
    torch.cuda.synchronize()
    out = grad_rf.cpu().numpy(), grad_c.cpu().numpy() / 1000.0       # back to 1/(mm/s)
    dd.close()
    return out


CASES = [
    ("linear_hann", dict(), dict()),
    ("hamming_offset", dict(window=N.WINDOW_HAMMING, additional_offset=5), dict()),
    ("subdivided_interp", dict(interpolate_transmits=True),
     dict(num_scanlines=(64, 1), rx_scanline_subdivision=(2, 1))),
    # 96 depths over 40 mm: 11-22 samples per depth, so the delays of one
    # 64-depth chunk span more than the kernel's 256-sample ring (ring moved
    # to the chunk's largest target, the rest added to the flushed line).
    ("coarse_depths", dict(), dict(num_depths=96)),
    ("coarse_depths_interp", dict(interpolate_transmits=True),
     dict(num_depths=96, num_scanlines=(64, 1), rx_scanline_subdivision=(2, 1))),
    # 20,000-sample RF lines: far more than LDS holds per wave (the kernel
    # keeps a 256-sample ring of each line).
    ("long_lines", dict(num_samples=20000), dict(num_scanlines=(16, 1), num_depths=300)),
]


@pytest.mark.parametrize("name,ov,bfov", CASES, ids=[c[0] for c in CASES])
def test_adjoint_vs_oracle(device, name, ov, bfov):
    w = W.linear_2d(num_scatterers=80, **ov)
    if bfov:
        w = replace(w, bf=replace(w.bf, **bfov))
    g = w.geometry()
    rf = frames_rf(w, len(C_MPS))
    rng = np.random.default_rng(5)
    gi = rng.normal(size=(len(C_MPS), g.num_depths, g.num_rx)).astype(np.float32).astype(np.float64)
    got_rf, got_c = run_adjoint(w, rf, C_MPS, gi)
    a = oracle_args(w)
    c_mm = np.array(C_MPS) * 1000.0
    r32, c32, _ = O.das_adjoint(g, rf, w.fs_hz, a["f_number"], c_mm, a["window_type"], gi, a["window_parameter"],
                                a["interpolate_transmits"], a["additional_offset"], precision=32)
    rmx, cmx, _ = O.das_adjoint(g, rf, w.fs_hz, a["f_number"], c_mm, a["window_type"], gi, a["window_parameter"],
                                a["interpolate_transmits"], a["additional_offset"], precision=6432)
    r64, c64, _ = O.das_adjoint(g, rf, w.fs_hz, a["f_number"], c_mm, a["window_type"], gi, a["window_parameter"],
                                a["interpolate_transmits"], a["additional_offset"], precision=64)
    e32 = np.linalg.norm(got_rf - r32) / np.linalg.norm(r32)
    m32 = np.abs(got_rf - r32).max() / np.abs(r32).max()
    emx = np.linalg.norm(got_rf - rmx) / np.linalg.norm(rmx)
    e64 = np.linalg.norm(got_rf - r64) / np.linalg.norm(r64)
    ec32 = np.abs(got_c - c32) / np.abs(c32)
    ecmx = np.abs(got_c - cmx) / np.abs(cmx)
    ec64 = np.abs(got_c - c64) / np.abs(c64)
    print(f"{name}: dL/drf vs FP32 oracle normwise {e32:.2e} max {m32:.2e}, vs FP64 on FP32 decisions {emx:.2e}, "
          f"vs FP64 {e64:.2e}; dL/dc vs FP32 {ec32.max():.2e}, vs FP64 on FP32 decisions {ecmx.max():.2e}, "
          f"vs FP64 {ec64.max():.2e}")
    assert e32 <= 1e-5 and m32 <= 1e-5
    assert ec32.max() <= 1e-5
    # Single-precision delays (ulp 1.2e-4 samples at tau ~ 2000) against
    # double-precision values on the same decisions.
    assert emx <= 1e-3 and ecmx.max() <= 1e-4
    # Plain FP64 also resolves exact ties (window arguments at pitch
    # multiples, membership at d/2F = |dx|) its own way.
    assert e64 <= 1e-2


def test_adjoint_deterministic(device):
    w = W.linear_2d(num_scatterers=40)
    g = w.geometry()
    rf = frames_rf(w, 2)
    gi = np.random.default_rng(2).normal(size=(2, g.num_depths, g.num_rx))
    a1 = run_adjoint(w, rf, [1480.0, 1560.0], gi)
    a2 = run_adjoint(w, rf, [1480.0, 1560.0], gi)
    assert np.array_equal(a1[0], a2[0]) and np.array_equal(a1[1], a2[1])


def test_autograd_loss_vs_reverse_mode(device):
    """loss = ||DAS(rf; c) - target||^2, target at 1540 m/s: the autograd
    wrapper against the reverse modes of the oracle."""
    w = W.linear_2d(num_scatterers=120)
    g = w.geometry()
    n = 2
    rf = frames_rf(w, n, seed0=30)
    dd = make(w)
    x = torch.from_numpy(rf).cuda()
    with torch.no_grad():
        target = dd(x, torch.full((n,), 1540.0, dtype=torch.float64, device="cuda"))
    xr = x.clone().requires_grad_(True)
    c = torch.tensor([1490.0, 1575.0], dtype=torch.float64, device="cuda", requires_grad=True)
    y = dd(xr, c)
    diff = torch.where((y.abs() < 3e38) & (target.abs() < 3e38), y - target, torch.zeros_like(y))
    loss = (diff * diff).sum()
    loss.backward()
    torch.cuda.synchronize()
    gi = (2.0 * diff).detach().double().cpu().numpy()
    a = oracle_args(w)
    res = {}
    for p in (32, 6432, 64):
        r, cc, _ = O.das_adjoint(g, rf, w.fs_hz, a["f_number"], np.array([1490e3, 1575e3]), a["window_type"], gi,
                                 precision=p)
        res[p] = (np.linalg.norm(xr.grad.cpu().numpy() - r) / np.linalg.norm(r),
                  (np.abs(c.grad.cpu().numpy() - cc * 1000.0) / np.abs(cc * 1000.0)).max())
    print("autograd loss gradients vs reverse mode (dL/drf normwise, dL/dc): FP32 %.2e %.2e; "
          "FP64 on FP32 decisions %.2e %.2e; FP64 %.2e %.2e" % (res[32] + res[6432] + res[64]))
    assert res[32][0] <= 1e-5 and res[32][1] <= 1e-5
    assert res[6432][0] <= 1e-3 and res[6432][1] <= 1e-4
    assert res[64][0] <= 1e-2

    dd.close()


def test_autograd_vs_plain_torch(device):
    """The same loss in plain PyTorch autograd (float32 delays from millimetre
    tables, float64 membership): with a rectangular window the two differ
    only by single-precision rounding of the delays and rare membership ties
    (window-table ties at pitch multiples would dominate dL/dc otherwise)."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).parent))
    from torch_das_reference import PlainTorchDAS

    w = W.linear_2d(num_scatterers=120, window=N.WINDOW_RECTANGULAR)
    g = w.geometry()
    n = 2
    rf = frames_rf(w, n, seed0=30)
    dd = make(w)
    x = torch.from_numpy(rf).cuda()
    with torch.no_grad():
        target = dd(x, torch.full((n,), 1540.0, dtype=torch.float64, device="cuda"))
    xr = x.clone().requires_grad_(True)
    c = torch.tensor([1490.0, 1575.0], dtype=torch.float64, device="cuda", requires_grad=True)
    y = dd(xr, c)
    ok = (y.abs() < 3e38) & (target.abs() < 3e38)
    diff = torch.where(ok, y - target, torch.zeros_like(y))
    (diff * diff).sum().backward()
    a = oracle_args(w)
    plain = PlainTorchDAS(g, w.fs_hz, a["f_number"], a["window_type"], a["window_parameter"], rf.shape[2],
                          rf.shape[3])
    for f in range(n):
        xf = x[f].clone().requires_grad_(True)
        cf = torch.tensor(float(c[f]), dtype=torch.float64, device="cuda", requires_grad=True)
        yp = plain(xf, cf)
        dp = torch.where(ok[f], yp - target[f], torch.zeros_like(yp))
        (dp * dp).sum().backward()
        m = ok[f]
        ey = (torch.linalg.norm((yp - y[f])[m]) / torch.linalg.norm(y[f][m])).item()
        eg = (torch.linalg.norm(xf.grad - xr.grad[f]) / torch.linalg.norm(xr.grad[f])).item()
        ec = abs(cf.grad.item() - c.grad[f].item()) / abs(c.grad[f].item())
        print(f"frame {f}: plain torch vs kernels: image {ey:.2e}, dL/drf {eg:.2e}, dL/dc {ec:.2e}")
        # dL/dc of this loss is a sum of terms of both signs: its relative
        # difference amplifies the float32 rounding of the plain form.
        assert ey <= 5e-4 and eg <= 1e-3 and ec <= 3e-2
    dd.close()


def test_rejects_3d_and_bad_sizes(device):
    from supra_amd.differentiable import DifferentiableDAS
    w3 = W.matrix_3d_small()
    with pytest.raises(ValueError):
        DifferentiableDAS(w3.geometry(), w3.fs_hz)
    w = W.linear_2d(num_scatterers=5)
    dd = make(w)
    x = torch.zeros((1, 128, 64, 2080), device="cuda")
    with pytest.raises(N.SupraError):
        dd.forward_raw(x[:, :100].contiguous(), torch.tensor([1540.0], dtype=torch.float64, device="cuda"))
    dd.close()
