"""Benchmark of the receive beamformer on one MI355X.

`python bench.py [--gpus 1] [--steps K] [--warmup W]` prints ONE JSON line.

Headline workload (value): 3D delay-and-sum of the 32x32 matrix probe —
64x64 scanlines x 1000 depths from int16 RF [4096 tx][1024 ch][1000 s]
(8.39 GB resident in HBM); one step = one volume.  The 2D linear-array
workload (128 x 2080 image from 64-channel int16 RF) is reported under
"extra", beamformed 32 frames per launch so that the 1.1 GB of RF exceeds the
256 MB Infinity Cache.

roofline.achieved = algorithmic bytes per launch (RF read once + image written
once) / average kernel duration from HIP events on the launch stream.
cpu_baseline = the host FP32 oracle (same algorithm, C++, all cores) on a
bounded sample of the same volume.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
METRIC = "beamformed samples/s & % HBM peak (DAS 2D/3D); MV pixels/s; fwd+adj ms"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def host_cpu() -> dict:
    """The host cores this process may run on (the CPU baseline uses all of
    them), the machine's logical CPU count, its CPU model and the cgroup CPU
    quota if one is set."""
    info = {"threads": max(1, len(os.sched_getaffinity(0))), "nproc": os.cpu_count() or 1, "model": None,
            "cgroup_cpu_quota": None}
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                info["model"] = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        quota, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if quota != "max":
            info["cgroup_cpu_quota"] = float(quota) / float(period)
    except (OSError, ValueError):
        pass
    return info


def normwise(a, ref) -> float:
    """max |a - ref| / max |ref| over samples that saturate in neither (the
    single-precision 0/0 of an element-aligned depth-0 pixel clamps to
    -FLT_MAX; double precision does not see that element at depth 0)."""
    a = np.asarray(a, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    m = (np.abs(ref) < 1e30) & (np.abs(a) < 1e30)
    return float(np.abs(a[m] - ref[m]).max() / np.abs(ref[m]).max())


def das_bytes(w, frames: int = 1) -> int:
    g = w.geometry()
    rf_bytes = w.num_tx * w.num_channels * w.num_samples * 2
    return frames * (rf_bytes + g.num_rx * g.num_depths * 4)


def time_launches(fn, steps: int, warmup: int, torch):
    """Warm up, then time `steps` launches with per-launch HIP events on the
    current stream; returns (wall ms per step, mean event ms per launch)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        starts[i].record()
        fn()
        ends[i].record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1000.0 / steps
    ev = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    return wall, ev


def bench_3d(args, torch):
    from supra_amd import workloads as W
    from supra_amd.beamformer import RxBeamformer, USRawData

    w = W.matrix_3d()
    g = w.geometry()
    log(f"[bench] synthesizing {w.name} RF {w.rf_shape} int16 on the GPU")
    rf = torch.empty(w.rf_shape, dtype=torch.int16, device="cuda")
    w.rf_device(rf.data_ptr(), stream=torch.cuda.current_stream().cuda_stream)
    bf = RxBeamformer(g)
    raw = USRawData(rf, w.fs_hz, g)
    out = torch.empty((g.num_depths, g.num_rx), dtype=torch.float32, device="cuda")
    kw = w.das_kwargs()

    def step():
        bf.perform_rx_beamforming(0, raw, kw["f_number"], kw["speed_of_sound_mm_per_s"], kw["window_type"],
                                  kw["window_parameter"], kw["interpolate_transmits"], kw["additional_offset"],
                                  out=out)

    log(f"[bench] timing {args.steps} volumes after {args.warmup} warm-up")
    wall_ms, ev_ms = time_launches(step, args.steps, args.warmup, torch)
    samples = g.num_rx * g.num_depths
    return dict(workload=w, geom=g, rf=rf, out=out, wall_ms=wall_ms, ev_ms=ev_ms, samples=samples,
                bytes=das_bytes(w), bf=bf)


def bench_2d(args, torch, frames: int = 32):
    from supra_amd import workloads as W
    from supra_amd.beamformer import RxBeamformer, USRawData

    w = W.linear_2d()
    g = w.geometry()
    log(f"[bench] synthesizing {frames} frames of {w.name} RF on the GPU")
    rf = torch.empty((frames,) + w.rf_shape, dtype=torch.int16, device="cuda")
    for f in range(frames):
        wf = W.linear_2d(scatter_seed=100 + f, noise_seed=200 + f)
        wf.rf_device(rf[f].data_ptr(), stream=torch.cuda.current_stream().cuda_stream)
    bf = RxBeamformer(g)
    raw = USRawData(rf, w.fs_hz, g)
    out = torch.empty((frames, g.num_depths, g.num_rx), dtype=torch.float32, device="cuda")
    kw = w.das_kwargs()

    def step():
        bf.perform_rx_beamforming(0, raw, kw["f_number"], kw["speed_of_sound_mm_per_s"], kw["window_type"],
                                  kw["window_parameter"], kw["interpolate_transmits"], kw["additional_offset"],
                                  out=out)

    wall_ms, ev_ms = time_launches(step, max(5, args.steps), 2, torch)
    per_frame_ms = ev_ms / frames
    samples = g.num_rx * g.num_depths
    gbs = das_bytes(w, frames) / (ev_ms * 1e-3) / 1e9
    # one frame per call, as BeamformingNode calls its engine
    raw1 = USRawData(rf[0], w.fs_hz, g)
    out1 = out[0]

    def step1():
        bf.perform_rx_beamforming(0, raw1, kw["f_number"], kw["speed_of_sound_mm_per_s"], kw["window_type"],
                                  kw["window_parameter"], kw["interpolate_transmits"], kw["additional_offset"],
                                  out=out1)

    _, ev1_ms = time_launches(step1, max(5, args.steps), 2, torch)
    bf.close()
    res = {
        "workload": f"{w.name}: 128x2080 image, int16 RF 128x64x2080, {frames} frames per launch",
        "samples_per_s": samples / (per_frame_ms * 1e-3),
        "us_per_frame": per_frame_ms * 1e3,
        "achieved_GBs": gbs,
        "hbm_frac": gbs / HBM_PEAK_GBS,
        "single_frame_call_us": ev1_ms * 1e3,
        "single_frame_hbm_frac": das_bytes(w, 1) / (ev1_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_2d(w, g, rf, out, kw, args.cpu_budget)
        res["gpu_over_cpu"] = res["samples_per_s"] / res["cpu_baseline"]["value"]
    return res


def cpu_baseline_2d(w, g, rf, out, kw, budget_s: float):
    """Host FP32 oracle on whole frames of the same 2D batch, one frame after
    another (cycling through the batch) until budget_s has passed; each
    frame's first pass is checked against the GPU image."""
    from oracle import oracle as O

    cpu = host_cpu()
    threads = cpu["threads"]
    rf_host = rf.cpu().numpy()
    out_host = out.cpu().numpy()
    n = rf_host.shape[0]
    done, exact, elapsed = 0, True, 0.0
    while done == 0 or elapsed < budget_s:
        f = done % n
        t0 = time.perf_counter()
        img = O.das(g, rf_host[f], w.fs_hz, threads=threads, **kw)
        elapsed += time.perf_counter() - t0
        if done < n:
            exact = exact and bool(np.array_equal(img, out_host[f]))
        done += 1
    ref64 = O.das(g, rf_host[0], w.fs_hz, precision=64, threads=threads, **kw)
    return {"value": done * g.num_rx * g.num_depths / elapsed, "unit": "samples/s", "cores": threads,
            "kind": "port", "sample": f"{done} whole frames, cycling through the {n}-frame batch "
                                      f"({done * g.num_rx * g.num_depths} samples, {elapsed:.1f} s)",
            "host": cpu, "gpu_matches_sample_bit_exact": exact,
            "gpu_normwise_err_vs_fp64": normwise(out_host[0], ref64)}


def bench_diff(args, torch, frames: int = 16):
    """Differentiable 2D DAS: loss = ||DAS(rf; c_f) - target_f||^2 over a
    batch of FP32 frames, c_f ~ U[1450, 1600] m/s (seed 4), targets
    beamformed at 1540 m/s.  One step = forward + loss gradient + adjoint
    (dL/drf and dL/dc for all frames).  Compared with the same loss written
    in plain PyTorch (gather + autograd), one frame at a time."""
    from supra_amd import workloads as W
    from supra_amd.differentiable import DifferentiableDAS
    sys.path.insert(0, str(ROOT / "tests"))
    from torch_das_reference import PlainTorchDAS

    w = W.linear_2d()
    g = w.geometry()
    log(f"[bench] synthesizing {frames} FP32 frames of {w.name} RF for the differentiable path")
    rf = torch.empty((frames,) + w.rf_shape, dtype=torch.float32, device="cuda")
    for f in range(frames):
        W.linear_2d(scatter_seed=300 + f, noise_seed=400 + f).rf_device(
            rf[f].data_ptr(), dtype=np.float32, stream=torch.cuda.current_stream().cuda_stream)
    c = torch.tensor(np.random.default_rng(4).uniform(1450.0, 1600.0, frames), dtype=torch.float64, device="cuda")
    kw = w.das_kwargs()
    dd = DifferentiableDAS(g, w.fs_hz, kw["f_number"], kw["window_type"], kw["window_parameter"])
    target, _ = dd.forward_raw(rf, torch.full((frames,), 1540.0, dtype=torch.float64, device="cuda"),
                               need_norm=False)
    tmask = target.abs() < 3e38

    def step():
        y, norm = dd.forward_raw(rf, c)
        diff = torch.where(tmask & (y.abs() < 3e38), y - target, 0.0)
        return dd.adjoint_raw(rf, c, norm, 2.0 * diff)

    log(f"[bench] timing fused forward+adjoint, {frames} frames per step")
    wall_ms, ev_ms = time_launches(step, max(5, args.steps), 2, torch)

    plain = PlainTorchDAS(g, w.fs_hz, kw["f_number"], kw["window_type"], kw["window_parameter"],
                          w.num_channels, w.num_samples)

    def plain_step():
        x = rf.detach().requires_grad_(True)
        cc = c.detach().clone().requires_grad_(True)
        for f in range(frames):
            yp = plain(x[f], cc[f])
            d = torch.where(tmask[f], yp - target[f], 0.0)
            (d * d).sum().backward()
        return x.grad, cc.grad

    log("[bench] timing the plain per-frame PyTorch form")
    p_wall, p_ev = time_launches(plain_step, 3, 1, torch)
    b = 2 * rf.numel() * 4 + 3 * frames * g.num_rx * g.num_depths * 4
    res = {
        "workload": f"linear_2d geometry, FP32 RF {frames}x128x64x2080, c ~ U[1450,1600] m/s, "
                    "loss ||DAS - DAS@1540||^2, dL/drf + dL/dc",
        "fwd_adj_ms": ev_ms,
        "plain_torch_per_frame_ms": p_ev,
        "speedup_vs_plain": p_ev / ev_ms,
        "algorithmic_bytes": b,
        "achieved_GBs": b / (ev_ms * 1e-3) / 1e9,
        "hbm_frac": b / (ev_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
    }
    if not args.no_cpu:
        log("[bench] gradient errors against the FP64 host reverse mode")
        res["gradient_errors"] = diff_errors(dd, rf, c, target, tmask, g, w, torch)
    dd.close()
    return res


def bench_bmode(args, torch, frames: int = 32):
    """After the beamformer, per linear_2d frame (2080 x 128 float image):
    envelope + log compression to uint8 of all frames in one launch
    (HilbertFirEnvelope 65 taps -> LogCompressor), then scan conversion of
    all frames in one launch onto the default grid (pixel pitch = sample
    distance, 1982 x 2080 pixels), uint8 -> uint8.  Algorithmic bytes: image
    read + B-mode write for the first stage; for the second, the tables once
    per launch (12 B/pixel) plus per frame the B-mode image and the output."""
    from supra_amd import workloads as W
    from supra_amd.beamformer import RxBeamformer, USRawData
    from supra_amd.postprocess import envelope_log_compress
    from supra_amd.scanconv import ScanConverter

    w = W.linear_2d()
    g = w.geometry()
    rf = torch.empty((frames,) + w.rf_shape, dtype=torch.int16, device="cuda")
    for f in range(frames):
        W.linear_2d(scatter_seed=100 + f, noise_seed=200 + f).rf_device(
            rf[f].data_ptr(), stream=torch.cuda.current_stream().cuda_stream)
    bf = RxBeamformer(g)
    kw = w.das_kwargs()
    img = bf.perform_rx_beamforming(0, USRawData(rf, w.fs_hz, g), kw["f_number"], kw["speed_of_sound_mm_per_s"],
                                    kw["window_type"], kw["window_parameter"], kw["interpolate_transmits"],
                                    kw["additional_offset"]).data
    bmode = torch.empty((frames, g.num_depths, g.num_rx), dtype=torch.uint8, device="cuda")

    def env_step():
        envelope_log_compress(img, 65, 80.0, 1.0, 20000.0, out_dtype=torch.uint8, out=bmode)

    log(f"[bench] timing envelope + log compression, {frames} frames")
    _, env_ms = time_launches(env_step, max(5, args.steps), 2, torch)
    sc = ScanConverter()
    t0 = time.perf_counter()
    sc.update_internals(g)
    tables_s = time.perf_counter() - t0
    out = torch.empty((frames,) + sc._shape(), dtype=torch.uint8, device="cuda")

    def sc_step():
        sc.convert(bmode, torch.uint8, out=out)

    log(f"[bench] timing scan conversion of {frames} frames onto {sc.image_size[:2]} pixels")
    _, sc_ms = time_launches(sc_step, max(5, args.steps), 2, torch)
    px = sc.num_pixels
    img_px = g.num_depths * g.num_rx
    env_bytes = frames * img_px * (4 + 1)
    sc_bytes = px * 12 + frames * (img_px + px)
    bf.close()
    sc.close()
    return {
        "workload": f"linear_2d images 2080x128 float, {frames} frames; scan conversion to "
                    f"{sc.image_size[0]}x{sc.image_size[1]} pixels, uint8",
        "envelope_log_us_per_frame": env_ms * 1e3 / frames,
        "envelope_log_hbm_frac": env_bytes / (env_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "scanconv_us_per_frame": sc_ms * 1e3 / frames,
        "scanconv_achieved_GBs": sc_bytes / (sc_ms * 1e-3) / 1e9,
        "scanconv_hbm_frac": sc_bytes / (sc_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "scanconv_tables_s": tables_s,
    }


def diff_errors(dd, rf, c, target, tmask, g, w, torch):
    """The step's dL/drf (normwise) and dL/dc (max relative over frames)
    against the FP64 host reverse mode, on all pixels and with the pixels
    where single and double precision take different discrete decisions
    (window-table ties, membership, floor of the delay) left out of the loss."""
    from oracle import oracle as O
    from supra_amd import _native as N

    threads = host_cpu()["threads"]
    y, norm = dd.forward_raw(rf, c)
    gi = torch.where(tmask & (y.abs() < 3e38), 2.0 * (y - target), 0.0)
    c_np = c.cpu().numpy()
    rf_host = rf.cpu().numpy()
    flags = np.stack([O.das_decision_diff(g, w.num_samples, w.fs_hz, 1.0, cf * 1000.0, N.WINDOW_HANN,
                                          threads=threads) for cf in c_np])
    out = {"tie_pixels_per_frame": float((flags != 0).sum() / len(c_np)), "pixels_per_frame": int(flags[0].size)}
    for label, keep in (("all_pixels", None), ("tie_pixels_out", flags == 0)):
        gi_l = gi if keep is None else gi * torch.from_numpy(keep).to(gi.device)
        g_rf, g_c = dd.adjoint_raw(rf, c, norm, gi_l)
        r64, c64, _ = O.das_adjoint(g, rf_host, w.fs_hz, 1.0, c_np * 1000.0, N.WINDOW_HANN,
                                    gi_l.double().cpu().numpy(), precision=64, threads=threads)
        c64 = c64 * 1000.0
        g_rf = g_rf.cpu().numpy()
        g_c = g_c.cpu().numpy()
        out[label] = {"dL_drf_normwise_vs_fp64": float(np.linalg.norm(g_rf - r64) / np.linalg.norm(r64)),
                      "dL_dc_max_rel_vs_fp64": float((np.abs(g_c - c64) / np.abs(c64)).max())}
    return out


MV_L, MV_T, MV_P, MV_ALPHA = 32, 5, 1.5, 0.01


def mv_flops_per_pixel(n_ch: int, L: int, T: int) -> float:
    """The reference's algorithm per output sample: K full L x L outer
    products, the temporal sum of 2T+1 matrices, an L^3/3 factorisation, two
    triangular solves, the subarray sums and the weights."""
    K = n_ch - L + 1
    return 2 * K * L * L + (2 * T + 1) * L * L + L ** 3 / 3 + 2 * L * L + K * L + 6 * L


def bench_mv(args, torch):
    """Minimum variance on the 2D workload: RawDelay (F#1, Hann) of one int16
    frame, then L = 32, T = 5, loading trace/(100 L) over 128 x 2080 samples
    (266,240 32x32 systems) in one launch.  Compared with the same algorithm
    in plain PyTorch on the whole frame (gather, batched matmul,
    torch.linalg.solve)."""
    from supra_amd import _native as N
    from supra_amd import workloads as W
    from supra_amd.beamformer import (RawDelay, RxBeamformer, USRawData, mv_das_perform_rx_beamforming,
                                      mv_perform_rx_beamforming)
    sys.path.insert(0, str(ROOT / "tests"))
    from torch_das_reference import PlainTorchDAS
    from torch_mv_reference import plain_mv

    w = W.linear_2d()
    g = w.geometry()
    rf = torch.empty(w.rf_shape, dtype=torch.int16, device="cuda")
    w.rf_device(rf.data_ptr(), stream=torch.cuda.current_stream().cuda_stream)
    rd = RawDelay(g)
    raw = USRawData(rf, w.fs_hz, g)
    kw = w.das_kwargs()
    delayed = rd.perform_delay(raw, kw["f_number"], kw["speed_of_sound_mm_per_s"], N.WINDOW_HANN, 0.0)
    out = torch.empty((g.num_depths, g.num_rx), dtype=torch.float32, device="cuda")

    def mv_step():
        mv_perform_rx_beamforming(delayed, MV_L, MV_T, MV_P, False, MV_ALPHA, out=out)

    def delay_step():
        rd.perform_delay(raw, kw["f_number"], kw["speed_of_sound_mm_per_s"], N.WINDOW_HANN, 0.0, out=delayed.data)

    engine = RxBeamformer(g)

    def fused_step():
        mv_das_perform_rx_beamforming(engine, raw, kw["f_number"], kw["speed_of_sound_mm_per_s"], N.WINDOW_HANN,
                                      0.0, MV_L, MV_T, MV_P, False, MV_ALPHA, out=out)

    log("[bench] timing minimum variance, the pre-delay, and the fused RF -> image launch")
    _, mv_ms = time_launches(mv_step, max(5, args.steps), 2, torch)
    _, rd_ms = time_launches(delay_step, max(5, args.steps), 2, torch)
    _, fused_ms = time_launches(fused_step, max(5, args.steps), 2, torch)
    log("[bench] timing the plain PyTorch pre-delay and minimum variance")
    _, plain_ms = time_launches(lambda: plain_mv(delayed.data, MV_L, MV_T, MV_P, MV_ALPHA), 3, 1, torch)
    pdas = PlainTorchDAS(g, w.fs_hz, kw["f_number"], N.WINDOW_HANN, 0.0, w.num_channels, w.rf_shape[2])
    rf32 = rf.to(torch.float32)
    c_t = torch.tensor(kw["speed_of_sound_mm_per_s"] / 1000.0, device="cuda")
    _, plain_rd_ms = time_launches(lambda: pdas.delayed(rf32, c_t, w.num_channels), 3, 1, torch)
    pixels = g.num_rx * g.num_depths
    flops = mv_flops_per_pixel(w.num_channels, MV_L, MV_T) * pixels
    err = None
    if not args.no_cpu:
        log("[bench] minimum variance against the FP64 host oracle")
        err = mv_errors(delayed.data, out, MV_L, MV_T, MV_P, MV_ALPHA, mv_step, torch)
    rd.close()
    engine.close()
    return {
        "error_vs_fp64_oracle": err,
        "workload": "linear_2d frame -> RawDelay (F#1 Hann) -> MV L=32 T=5 loading trace/(100L), "
                    "128x2080 samples",
        "mv_pixels_per_s": pixels / (mv_ms * 1e-3),
        "mv_ms": mv_ms,
        "rawdelay_ms": rd_ms,
        "plain_torch_mv_ms": plain_ms,
        "speedup_vs_plain": plain_ms / mv_ms,
        "rf_to_image": {
            "fused_ms": fused_ms,
            "two_step_ms": rd_ms + mv_ms,
            "plain_torch_ms": plain_rd_ms + plain_ms,
            "plain_torch_rawdelay_ms": plain_rd_ms,
            "speedup_vs_plain": (plain_rd_ms + plain_ms) / fused_ms,
            "two_step_speedup_vs_plain": (plain_rd_ms + plain_ms) / (rd_ms + mv_ms),
        },
        "algorithmic_flop": flops,
        "achieved_TFLOPs": flops / (mv_ms * 1e-3) / 1e12,
        "fp32_frac": flops / (mv_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
    }


def mv_errors(delayed, out, L, T, p, alpha, mv_step, torch):
    """Per-sample relative error of the MV image against the FP64 host oracle
    (denominator floored at 1e-3 of the image maximum); samples whose
    temporal window holds a -FLT_MAX pre-delayed value are 0 in single
    precision, as in the reference, and are left out."""
    from oracle import oracle as O

    mv_step()
    torch.cuda.synchronize()
    x = delayed.cpu().numpy()
    gpu = out.cpu().numpy().astype(np.float64)
    ref = O.mv(x, L, T, p, False, alpha, threads=host_cpu()["threads"])
    huge = (np.abs(x) > 1e19).any(axis=1)
    tainted = np.zeros_like(huge)
    for d in range(-T, T + 1):
        tainted |= np.roll(huge, d, axis=1)
    keep = ~tainted.T
    den = np.maximum(np.abs(ref[keep]), 1e-3 * np.abs(ref[keep]).max())
    e = np.abs(gpu[keep] - ref[keep]) / den
    return {"samples": int(keep.sum()), "max": float(e.max()), "p99": float(np.quantile(e, 0.99)),
            "median": float(np.median(e))}


def bench_stream(args, torch, num: int = 48):
    """Uneven batched stream: 48 seeded acquisitions (scanlines 64-256, depth
    samples 1024-4096, apertures 32-128 channels, fs 20/40 MHz), each by
    delay-and-sum or pre-delay + minimum variance (L = aperture/2, T = 5,
    loading trace/(100 L)), all in one graph launch of persistent kernels
    with autotuned tiles.  Latency = an acquisition's completion time after
    the launch starts (device timestamps).  Compared with the same
    acquisitions issued one call each on one stream."""
    from supra_amd import workloads as W
    from supra_amd import _native as N
    from supra_amd.batch import DAS, MV, Batch
    from supra_amd.beamformer import RxBeamformer, USRawData, mv_das_perform_rx_beamforming

    acqs = W.uneven_stream(num, seed=5)
    log(f"[bench] synthesizing the {num}-acquisition uneven stream on the GPU")
    items, b_das, f_mv, pixels = [], 0, 0.0, 0
    stream = torch.cuda.current_stream().cuda_stream
    for w, mode in acqs:
        g = w.geometry()
        rf = torch.empty(w.rf_shape, dtype=torch.int16, device="cuda")
        w.rf_device(rf.data_ptr(), stream=stream)
        out = torch.empty((g.num_depths, g.num_rx), dtype=torch.float32, device="cuda")
        it = dict(engine=RxBeamformer(g), rf=rf, out=out, fs_hz=w.fs_hz, mode=DAS if mode == "das" else MV,
                  window_type=N.WINDOW_HANN, sub_array_size=w.num_channels // 2, temporal_smoothing=5,
                  loading_factor=0.01, sub_array_scaling_power=1.5)
        items.append(it)
        pixels += g.num_rx * g.num_depths
        if mode == "das":
            b_das += das_bytes(w)
        else:
            f_mv += mv_flops_per_pixel(w.num_channels, w.num_channels // 2, 5) * g.num_rx * g.num_depths
    torch.cuda.synchronize()
    log("[bench] planning the batch (autotuner) and capturing the graph")
    batch = Batch(items, autotune=True)

    def step():
        batch.launch()

    steps = max(3, min(args.steps, 5))
    wall_ms, ev_ms = time_launches(step, steps, 2, torch)
    lat = []
    for _ in range(steps):
        batch.launch()
        l, _ = batch.latencies()
        lat.append(l)
    is_das = np.tile(np.array([m == "das" for _, m in acqs]), steps)
    lat = np.concatenate(lat) / 1e3

    def sequential():
        for it in items:
            raw = USRawData(it["rf"], it["fs_hz"], it["engine"].geometry)
            if it["mode"] == DAS:
                it["engine"].perform_rx_beamforming(N.BF_DELAY_AND_SUM, raw, 1.0, 1540e3, N.WINDOW_HANN, 0.0, False,
                                                    0, out=it["out"])
            else:
                mv_das_perform_rx_beamforming(it["engine"], raw, 1.0, 1540e3, N.WINDOW_HANN, 0.0,
                                              it["sub_array_size"], 5, 1.5, False, 0.01, out=it["out"])

    log("[bench] timing the same acquisitions one call each")
    _, seq_ms = time_launches(sequential, 2, 1, torch)
    config = batch.config()
    batch.close()
    return {
        "workload": f"{num} acquisitions (seed 5): {sum(m == 'das' for _, m in acqs)} DAS, "
                    f"{sum(m == 'mv' for _, m in acqs)} MV, {pixels} output samples",
        "frames_per_s": num / (ev_ms * 1e-3),
        "ms_per_launch": ev_ms,
        "latency_ms_p50": float(np.percentile(lat, 50)),
        "latency_ms_p99": float(np.percentile(lat, 99)),
        "das_latency_ms_p50": float(np.percentile(lat[is_das], 50)),
        "mv_latency_ms_p50": float(np.percentile(lat[~is_das], 50)),
        "sequential_calls_ms": seq_ms,
        "speedup_vs_sequential_calls": seq_ms / ev_ms,
        "algorithmic_das_bytes": b_das,
        "algorithmic_mv_flop": f_mv,
        "autotune": config,
    }


def cpu_baseline_3d(res, budget_s: float):
    """Host FP32 oracle on whole scanlines of the same volume, doubling the
    sample until it takes >= budget_s (or covers the volume)."""
    from oracle import oracle as O

    w, g = res["workload"], res["geom"]
    cpu = host_cpu()
    threads = cpu["threads"]
    log(f"[bench] copying RF to host for the CPU baseline ({threads} threads)")
    rf_host = res["rf"].cpu().numpy()
    kw = w.das_kwargs()
    rng = np.random.default_rng(11)
    n_lines = 8
    while True:
        lines = np.sort(rng.choice(g.num_rx, size=min(n_lines, g.num_rx), replace=False))
        pix = (np.arange(g.num_depths)[:, None] * g.num_rx + lines[None, :]).ravel().astype(np.uint32)
        t0 = time.perf_counter()
        vals = O.das(g, rf_host, w.fs_hz, pixels=pix, threads=threads, **kw)
        dt = time.perf_counter() - t0
        if dt >= budget_s or n_lines >= g.num_rx:
            break
        n_lines *= 2
    gpu_all = res["out"].cpu().numpy().ravel()
    exact = bool(np.array_equal(gpu_all[pix], vals))
    sub = np.sort(rng.choice(g.num_rx, size=min(64, g.num_rx), replace=False))
    pix64 = (np.arange(g.num_depths)[:, None] * g.num_rx + sub[None, :]).ravel().astype(np.uint32)
    ref64 = O.das(g, rf_host, w.fs_hz, pixels=pix64, precision=64, threads=threads, **kw)
    return {"value": pix.size / dt, "unit": "samples/s", "cores": threads, "kind": "port",
            "sample": f"{lines.size} of {g.num_rx} scanlines x {g.num_depths} depths of the same volume "
                      f"({pix.size} samples, {dt:.1f} s)",
            "host": cpu, "gpu_matches_sample_bit_exact": exact,
            "gpu_normwise_err_vs_fp64": normwise(gpu_all[pix64], ref64),
            "fp64_sample": f"{sub.size} scanlines x {g.num_depths} depths"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--frames-2d", type=int, default=32, help="frames per launch of the 2D DAS leg")
    ap.add_argument("--no-2d", action="store_true")
    ap.add_argument("--no-diff", action="store_true")
    ap.add_argument("--no-mv", action="store_true")
    ap.add_argument("--no-stream", action="store_true")
    ap.add_argument("--no-bmode", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain device copy of 1 GiB (read 1 GiB + write 1 GiB) so that "
                         "FETCH_SIZE/WRITE_SIZE counters of a profile can be scaled")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this beamformer runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is a design note, not part of this build)"}))
        return 2
    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        return 1

    if args.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    res = bench_3d(args, torch)
    algo_gbs = res["bytes"] / (res["ev_ms"] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(algo_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(algo_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "das3d_kernel", "avg_launch_ms": round(res["ev_ms"], 4),
                "algorithmic_bytes_per_launch": res["bytes"]}
    traffic_file = ROOT / "profiles" / "traffic.json"
    if traffic_file.exists():
        t = json.loads(traffic_file.read_text()).get("das3d_kernel/matrix_3d")
        if t:
            roofline["traffic"] = t["hbm_bytes_per_launch"]
    extra = {}
    if not args.no_2d:
        extra["linear_2d"] = bench_2d(args, torch, frames=args.frames_2d)
    if not args.no_diff:
        extra["diff_2d"] = bench_diff(args, torch)
    if not args.no_mv:
        extra["mv_2d"] = bench_mv(args, torch)
    if not args.no_stream:
        extra["uneven_stream"] = bench_stream(args, torch)
    if not args.no_bmode:
        extra["bmode_2d"] = bench_bmode(args, torch)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_3d(res, args.cpu_budget)
    g = res["geom"]
    value = res["samples"] / (res["ms_per_step"] * 1e-3) if "ms_per_step" in res else \
        res["samples"] / (res["wall_ms"] * 1e-3)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["wall_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded point scatterers + noise, int16 RF resident in HBM)",
        "config": {"workload": "matrix_3d: 3D DAS, 32x32 matrix probe, 64x64 scanlines x 1000 depths, "
                               "int16 RF 4096x1024x1000, F#1 separable Hann, fs 20 MHz",
                   "rf_bytes": res["workload"].num_tx * 1024 * 1000 * 2, "out": [g.num_depths, g.num_rx]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
