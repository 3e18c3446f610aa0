"""Pins the host oracle's pre-delay (RawDelay.cu) and minimum-variance
beamformer (RxBeamformerMV.cu): channel sums of the pre-delay equal DAS,
plane-wave and infinite-loading closed forms of MV, and numpy.linalg.solve
on the systems the oracle assembles."""
import numpy as np
import pytest

from oracle import oracle as O
from supra_amd import _native as N
from supra_amd import workloads as W


def test_rawdelay_sums_to_das():
    w = W.linear_2d(num_scatterers=60)
    g = w.geometry()
    rf = w.rf_host()
    for win in (N.WINDOW_HANN, N.WINDOW_RECTANGULAR):
        kw = dict(w.das_kwargs(), window_type=win)
        rd = O.rawdelay(g, rf, w.fs_hz, kw["f_number"], kw["speed_of_sound_mm_per_s"], win, precision=64)
        das = O.das(g, rf, w.fs_hz, precision=64, **kw)
        s = rd.sum(1).T
        fin = np.isfinite(das) & (np.abs(das) < 1e30)
        assert np.abs(s[fin] - das[fin]).max() <= 1e-12 * np.abs(das[fin]).max()
        assert np.all(rd[:, 63, :] == 0)            # only last-first = 63 channels are written


def test_rawdelay3d_sums_to_das3d():
    """3D pre-delay: summing the slots of each pixel gives the 3D DAS volume
    (same F-number and window, no offset); slot (ex - first.x) + (ey -
    first.y) * layout.x carries element (ex, ey), other slots stay 0."""
    w = W.matrix_3d_small(n_el=6, n_lines=5, n_depths=120, n_samples=700)
    g = w.geometry()
    rf = w.rf_host()
    for win in (N.WINDOW_HANN, N.WINDOW_RECTANGULAR):
        kw = dict(w.das_kwargs(), window_type=win)
        rd = O.rawdelay(g, rf, w.fs_hz, kw["f_number"], kw["speed_of_sound_mm_per_s"], win, precision=64)
        das = O.das(g, rf, w.fs_hz, precision=64, **kw)
        assert rd.shape == (g.num_rx, rf.shape[1], g.num_depths)
        s = rd.sum(1).T
        fin = np.isfinite(das) & (np.abs(das) < 1e30)
        assert np.abs(s[fin] - das[fin]).max() <= 1e-11 * np.abs(das[fin]).max()
        first = g.scanlines["tx"][:, 0]["first"]
        last = g.scanlines["tx"][:, 0]["last"]
        lx = g.element_layout[0]
        for sl in range(g.num_rx):
            used = {(ex - first[sl, 0]) + (ey - first[sl, 1]) * lx
                    for ex in range(first[sl, 0], last[sl, 0]) for ey in range(first[sl, 1], last[sl, 1])}
            unused = sorted(set(range(rf.shape[1])) - used)
            assert np.all(rd[sl, unused, :] == 0)
    rf_c = np.full(w.rf_shape, 10, dtype=np.int16)
    rd = O.rawdelay(g, rf_c, w.fs_hz, 1.0, 1.54e6, N.WINDOW_RECTANGULAR, precision=64)
    assert set(np.unique(np.round(rd[:, :, 60:100], 9))) <= {0.0, 10.0}
    assert (np.round(rd[:, :, 60:100], 9) == 10.0).any()


def test_rawdelay_constant_rf_channel_weights():
    """Constant RF k: channel j carries w_j * k * n / W, so with a rectangular
    window every in-aperture channel equals k * n / n = k."""
    w = W.linear_2d()
    g = w.geometry()
    rf = np.full(w.rf_shape, 10, dtype=np.int16)
    rd = O.rawdelay(g, rf, w.fs_hz, 1.0, 1.54e6, N.WINDOW_RECTANGULAR, precision=64)
    vals = np.unique(np.round(rd[:, :, 1200:1900], 9))
    assert set(vals) <= {0.0, 10.0}


def plane_wave(n_rx=4, n_ch=64, n_d=120, seed=0):
    rng = np.random.default_rng(seed)
    s = (rng.uniform(1.0, 3.0, (n_rx, 1, n_d)) * rng.choice([-1.0, 1.0], (n_rx, 1, n_d))).astype(np.float32)
    return np.ascontiguousarray(np.broadcast_to(s, (n_rx, n_ch, n_d))), s[:, 0, :].astype(np.float64)


@pytest.mark.parametrize("alpha,T,p", [(0.01, 5, 1.5), (1.0, 0, 1.5), (3.0, 3, 1.0)])
def test_mv_plane_wave_closed_form(alpha, T, p):
    """All channels equal s(t): R~ = S 1 1^T, w = 1 / (S (L + delta/S)), out = N_ch s(t) K^(1-p)."""
    x, s = plane_wave()
    n_ch, L = 64, 32
    K = n_ch - L + 1
    out = O.mv(x, L, T, p, False, alpha)
    expected = (n_ch * s * K ** (1 - p)).T
    assert np.allclose(out, expected, rtol=1e-9)


def random_apertures(n_rx=3, n_ch=64, n_d=60, seed=1):
    """Random data whose non-zero channels form a per-sample window [f, l]."""
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n_rx, n_ch, n_d)).astype(np.float32)
    f = rng.integers(0, 20, (n_rx, n_d))
    l = np.minimum(n_ch - 1, f + rng.integers(5, 64, (n_rx, n_d)))
    ch = np.arange(n_ch)[None, :, None]
    x[(ch < f[:, None, :]) | (ch > l[:, None, :])] = 0.0
    x[x == 0] = 0.0
    return x, f, l


def test_mv_infinite_loading_closed_form():
    """alpha -> inf: R~ + delta I ~ delta I, so y = sum_i X_i / (L_t K^p) with
    X_i = sum over active subarrays of x[f + k + i]."""
    x, f, l = random_apertures()
    n_rx, n_ch, n_d = x.shape
    L, p = 32, 1.5
    K = n_ch - L + 1
    out = O.mv(x, L, 2, p, False, 1e12)
    for s in range(n_rx):
        for t in range(n_d):
            lt = min(L, l[s, t] - f[s, t] + 1)
            kact = min(K, l[s, t] - f[s, t] - lt + 2)
            X = sum(x[s, f[s, t] + k: f[s, t] + k + lt, t].astype(np.float64) for k in range(kact))
            y = X.sum() / (lt * kact ** p)
            assert abs(out[t, s] - n_ch * y) <= 1e-6 * max(1.0, abs(n_ch * y)), (s, t)


@pytest.mark.parametrize("means", [False, True])
def test_mv_system_matches_numpy(means):
    """The assembled R~ + delta I equals the definition evaluated in numpy
    (with the per-sample subarray mean removed when means is on), and the
    oracle's weights solve it."""
    x, f, l = random_apertures(seed=3)
    L, T, alpha = 32, 3, 0.05
    for (s, t) in [(0, 10), (1, 0), (2, 59), (1, 31)]:
        _, A, wts, meta = O.mv(x, L, T, 1.5, means, alpha, debug=(s, t))
        first, last, lt, kact = meta
        n_d = x.shape[2]
        R = np.zeros((L, L))
        for tau in range(max(0, t - T), min(n_d - 1, t + T) + 1):
            f_, l_ = f[s, tau], l[s, tau]
            lt_ = min(L, l_ - f_ + 1)
            k_ = min(x.shape[1] - L + 1, l_ - f_ - lt_ + 2)
            subs = np.stack([x[s, f_ + k: f_ + k + lt_, tau].astype(np.float64) for k in range(k_)])
            if means:
                subs = subs - subs.mean(axis=0)
            for v in subs:
                R[:lt_, :lt_] += np.outer(v, v)
        Rt = R[:lt, :lt]
        Rt = Rt + alpha * np.trace(Rt) / lt * np.eye(lt)
        assert np.allclose(A, Rt, rtol=1e-12, atol=1e-12)
        assert np.allclose(np.linalg.solve(A, np.ones(lt)), wts, rtol=1e-10)
        assert (first, last) == (f[s, t], l[s, t])


@pytest.mark.parametrize("alpha,T,p", [(0.01, 5, 1.5), (1.0, 0, 1.0), (2.0, 3, 1.5)])
def test_mv_means_channel_ramp_closed_form(alpha, T, p):
    """Mean removal on x[c, t] = a(t) + b(t) c: every subarray minus the
    subarray mean is b (k - (K-1)/2) 1, so R~ = S 1 1^T (S > 0) and, like the
    plane wave, w = 1 / (S (L + alpha)); the output is the mean-free
    beamformer of the subarray sums: y = K^(1-p) (a + b (f + (K-1)/2 +
    (L-1)/2)) with f = 0, times N_ch — independent of alpha and T.  A sign
    slip in the mean removal or a mean taken over the wrong index would
    leave a rank-2 R~ and break it."""
    rng = np.random.default_rng(12)
    n_rx, n_ch, n_d, L = 3, 64, 90, 32
    K = n_ch - L + 1
    # integers and quarters: the ramp is exact in float32
    a = rng.integers(200, 300, (n_rx, n_d)).astype(np.float64)
    b = rng.integers(4, 9, (n_rx, n_d)) / 4.0 * rng.choice([-1.0, 1.0], (n_rx, n_d))
    x = (a[:, None, :] + b[:, None, :] * np.arange(n_ch)[None, :, None]).astype(np.float32)
    assert np.array_equal(x.astype(np.float64), a[:, None, :] + b[:, None, :] * np.arange(n_ch)[None, :, None])
    out = O.mv(x, L, T, p, True, alpha)
    expected = (n_ch * K ** (1 - p) * (a + b * ((K - 1) / 2 + (L - 1) / 2))).T
    assert np.allclose(out, expected, rtol=1e-9), np.abs(out / expected - 1).max()


def test_mv_on_predelayed_frame_is_finite():
    w = W.linear_2d(num_scatterers=50)
    g = w.geometry()
    rf = w.rf_host()
    rd = O.rawdelay(g, rf, w.fs_hz, 1.0, 1.54e6, N.WINDOW_HANN)
    out = O.mv(rd[:8], 32, 5, 1.5, False, 0.01)
    assert np.isfinite(out).all() and np.abs(out).max() > 0


def test_mvpcg_full_iterations_equal_the_direct_solve():
    """Conjugate gradients end in n steps: with L_t iterations and no
    threshold the PCG oracle is the LU oracle (FP64)."""
    x, _, _ = random_apertures(seed=3)
    for L, T in ((16, 2), (32, 0)):
        a = O.mv(x, L, T, 1.5, False, 1.0)
        b = O.mvpcg(x, L, T, max_iterations=L, threshold=1e-30)
        assert np.abs(a - b).max() <= 1e-8 * np.abs(a).max()


def test_mvpcg_error_a_norm_never_grows():
    """The A-norm of the error of (preconditioned) conjugate gradients does
    not grow with the iteration count; w_k from k = 1..n iterations of one
    sample, w* from numpy.linalg.solve on the oracle's own system."""
    x, _, _ = random_apertures(seed=4)
    for (s, t) in ((0, 20), (2, 41)):
        _, A, w_star_o, meta = O.mvpcg(x, 16, 2, max_iterations=16, threshold=1e-30, debug=(s, t))
        w_star = np.linalg.solve(A, np.ones(len(A)))
        errs = []
        for k in range(1, len(A) + 1):
            _, _, wk, mk = O.mvpcg(x, 16, 2, max_iterations=k, threshold=1e-30, debug=(s, t))
            assert mk[4] <= k
            e = wk - w_star
            errs.append(e @ A @ e)
            if mk[4] < k:          # converged (r.M^-1 r <= 1e-30) before the budget
                break
        errs = np.array(errs)
        assert np.all(np.diff(errs) <= 1e-9 * errs[0]), errs
        assert errs[-1] <= 1e-16 * errs[0]


@pytest.mark.parametrize("T,p", [(3, 1.5), (0, 1.0)])
def test_mvpcg_plane_wave_one_iteration(T, p):
    """All channels equal: 1 is an eigenvector of R~ + delta I and of its
    diagonal, so one PCG step solves the system and the output is the MV
    closed form N_ch s(t) K^(1-p)."""
    x, s = plane_wave()
    n_ch, L = 64, 32
    K = n_ch - L + 1
    out = O.mvpcg(x, L, T, max_iterations=1, threshold=1e-30, scaling_power=p)
    assert np.allclose(out, (n_ch * s * K ** (1 - p)).T, rtol=1e-9)


def test_mvpcg_threshold_below_float_range_runs_no_iteration():
    """The reference hands its kernel (float)threshold: 10^-50 becomes 0 and
    the loop (residual 2 * threshold > threshold) never starts, so w = 1 and
    y = sum_i X_i / (L_t K^p), the infinite-loading closed form."""
    x, f, l = random_apertures(seed=5)
    n_rx, n_ch, n_d = x.shape
    L, p = 32, 1.5
    K = n_ch - L + 1
    out = O.mvpcg(x, L, 2, threshold=1e-50, scaling_power=p)
    _, _, _, meta = O.mvpcg(x, L, 2, threshold=1e-50, scaling_power=p, debug=(1, 7))
    assert meta[4] == 0
    for s in range(n_rx):
        for t in range(n_d):
            lt = min(L, l[s, t] - f[s, t] + 1)
            kact = min(K, l[s, t] - f[s, t] - lt + 2)
            X = sum(x[s, f[s, t] + k: f[s, t] + k + lt, t].astype(np.float64) for k in range(kact))
            y = X.sum() / (lt * kact ** p)
            assert abs(out[t, s] - n_ch * y) <= 1e-9 * max(1.0, abs(n_ch * y)), (s, t)
