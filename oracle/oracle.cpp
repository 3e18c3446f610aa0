// ============================================================================
// oracle.cpp — CPU restatement of SUPRA's receive beamformer.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke test
// of __graft_entry__ and the cpu_baseline leg of bench.py call it to produce
// the expected output of the HIP library on the same inputs.  Nothing in the
// product (supra_amd/) links, loads or calls it.
//
// Two precisions of the same algorithm:
//   * FP32: the reference's single-precision arithmetic restated literally —
//     same operand order, IEEE rounding, no FMA contraction (built with
//     -ffp-contract=off), fmaxf/fminf clamps that keep the non-NaN operand,
//     CUDA's saturating float->uint32 conversion where the reference relies on
//     it.  The HIP kernels are expected to match it bit for bit.
//   * FP64: the same intended math in double precision end to end (geometry
//     conversion, window table, delays, interpolation, sums): the accuracy
//     yardstick the FP32 results are reported against.
//
// Each function cites the reference lines it follows (paths relative to the
// SUPRA source tree).  What pins each function is listed in its header
// comment; a function without such a list would be "parity unpinned".
// ============================================================================
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "supra_bf.h"

namespace {

constexpr int kEntries = 64;

// Static partition of [0, n) over worker threads.
template <typename Fn>
void parallel_for(int64_t n, int nthreads, Fn fn, int64_t min_parallel = 64) {
    if (nthreads <= 1 || n < min_parallel) {
        for (int64_t i = 0; i < n; i++) fn(i);
        return;
    }
    int64_t per = (n + nthreads - 1) / nthreads;
    std::vector<std::thread> pool;
    for (int w = 0; w < nthreads; w++) {
        int64_t lo = w * per, hi = std::min<int64_t>(n, lo + per);
        if (lo >= hi) break;
        pool.emplace_back([=]() {
            for (int64_t i = lo; i < hi; i++) fn(i);
        });
    }
    for (auto& t : pool) t.join();
}

// --- scalar helpers with the reference's semantics, per precision ------------
inline float vmax(float a, float b) { return std::fmax(a, b); }
inline float vmin(float a, float b) { return std::fmin(a, b); }
inline double vmax(double a, double b) { return std::fmax(a, b); }
inline double vmin(double a, double b) { return std::fmin(a, b); }
inline float vsqrt(float a) { return std::sqrt(a); }
inline double vsqrt(double a) { return std::sqrt(a); }
inline float vfloor(float a) { return std::floor(a); }
inline double vfloor(double a) { return std::floor(a); }
inline float vround(float a) { return std::round(a); }
inline double vround(double a) { return std::round(a); }
inline float vcos(float a) { return std::cos(a); }
inline double vcos(double a) { return std::cos(a); }

// WindowFunction::windowFunction (src/SupraLib/Beamformer/WindowFunction.h:121-138)
// + normalisation (WindowFunction.cpp:18-47).
// Pinned by: tests/test_oracle_window.py — closed-form window values computed
// in numpy float64, max == 1 after normalisation, end points of Hann/Hamming,
// Gauss symmetry; golden tables in tests/golden/window_luts.json.
template <typename T>
void window_lut(int type, float p_in, T* lut) {
    const T p = static_cast<T>(p_in);
    T vmax_ = std::numeric_limits<T>::min();
    for (int i = 0; i < kEntries; i++) {
        T r = (static_cast<T>(i) / static_cast<T>(kEntries - 1)) * 2 - 1;
        T v;
        const T two_pi = 2 * static_cast<T>(M_PI);
        switch (type) {
        case SUPRA_WINDOW_RECTANGULAR: v = 1; break;
        case SUPRA_WINDOW_HANN: v = (1 - p) * (T(0.5) - T(0.5) * vcos(two_pi * ((r + 1) * T(0.5)))) + p; break;
        case SUPRA_WINDOW_HAMMING: v = (1 - p) * (T(0.54) - T(0.46) * vcos(two_pi * ((r + 1) * T(0.5)))) + p; break;
        case SUPRA_WINDOW_GAUSS: {
            T q = r / p;
            v = static_cast<T>(1.0 / (p * std::sqrt(2.0 * M_PI)) * std::exp((-1.0 / 2.0) * q * q));
            break;
        }
        default: v = 0;
        }
        lut[i] = v;
        vmax_ = std::max(vmax_, v);
    }
    for (int i = 0; i < kEntries; i++) lut[i] /= vmax_;
}

// WindowFunctionGpu::get (WindowFunction.h:58-64)
template <typename T>
inline T window_get(const T* lut, T rel) {
    T cl = vmin(vmax(rel, T(-1)), T(1));
    T idx = vround(T(31.5) * (cl + T(1)));
    return lut[static_cast<uint32_t>(idx)];
}

template <typename T>
inline T clamp_out(T x, int out_dtype) {
    if (out_dtype == SUPRA_INT16) {
        // clampCast<int16_t> (utilities/cudaUtility.h:111-147): truncation.
        return static_cast<T>(static_cast<int16_t>(vmin(vmax(x, T(-32767)), T(32767))));
    }
    return vmin(vmax(x, T(-FLT_MAX)), T(FLT_MAX));
}

// Transmit selection (RxBeamformerCuda.cu:274-295): running max held in float
// and compared against the double weights, ties keep the lowest index.
int select_tx(const supra_scanline_rx& s) {
    int best = 0;
    float best_w = static_cast<float>(s.tx_weights[0]);
    for (int k = 1; k < 4; k++) {
        if (s.tx_weights[k] > best_w) {
            best_w = static_cast<float>(s.tx_weights[k]);
            best = k;
        }
    }
    return best;
}

// dt-space geometry of RxBeamformerCuda::convertToDtSpace (RxBeamformerCuda.cu:58-107),
// as the kernels read it: float tables for FP32, doubles for FP64.
template <typename T>
struct DtGeometry {
    std::vector<T> depth, ex, ey;
    std::vector<T> sx, sz, dirx, diry, dirz, medx, medz;
    std::vector<T> t0;   // [scanline*4 + k]
};

template <typename T>
DtGeometry<T> to_dt_space(const supra_rx_geometry& g, double dt, double c) {
    DtGeometry<T> o;
    const double factor = 1 / 1.0 / (c * dt);
    const double factor_time = 1 / 1.0 / dt;
    o.depth.resize(g.num_depths);
    for (uint32_t i = 0; i < g.num_depths; i++) o.depth[i] = static_cast<T>(g.depths[i] * factor);
    o.ex.resize(g.num_elements);
    o.ey.resize(g.num_elements);
    for (uint32_t i = 0; i < g.num_elements; i++) {
        o.ex[i] = static_cast<T>(g.element_x[i] * factor);
        o.ey[i] = static_cast<T>(g.element_y[i] * factor);
    }
    const uint32_t n = g.num_rx_scanlines;
    o.sx.resize(n); o.sz.resize(n); o.dirx.resize(n); o.diry.resize(n); o.dirz.resize(n);
    o.medx.resize(n); o.medz.resize(n); o.t0.resize(4 * n);
    for (uint32_t i = 0; i < n; i++) {
        const supra_scanline_rx& s = g.scanlines[i];
        o.sx[i] = static_cast<T>(s.position[0] * factor);
        o.sz[i] = static_cast<T>(s.position[2] * factor);
        o.dirx[i] = static_cast<T>(s.direction[0]);
        o.diry[i] = static_cast<T>(s.direction[1]);
        o.dirz[i] = static_cast<T>(s.direction[2]);
        o.medx[i] = static_cast<T>(s.max_element_distance[0] * factor);
        o.medz[i] = static_cast<T>(s.max_element_distance[1] * factor);
        for (int k = 0; k < 4; k++) o.t0[4 * i + k] = static_cast<T>(s.tx[k].initial_delay * factor_time);
    }
    return o;
}

struct RfView {
    const void* data;
    int dtype;
    uint32_t n_tx, n_ch, n_s;
    inline double at(uint64_t idx) const {
        return dtype == SUPRA_INT16 ? static_cast<const int16_t*>(data)[idx] : static_cast<const float*>(data)[idx];
    }
};

template <typename T>
inline T rf_at(const RfView& rf, uint64_t idx) {
    if (rf.dtype == SUPRA_INT16) return static_cast<T>(static_cast<const int16_t*>(rf.data)[idx]);
    return static_cast<T>(static_cast<const float*>(rf.data)[idx]);
}

struct DasCall {
    const supra_rx_geometry* g;
    RfView rf;
    double fs, F, c;
    int window;
    float wparam;
    int interp_tx;
    int32_t offset;
    int kind;        // sample beamformer: SUPRA_BF_DELAY_AND_SUM, _DELAY_AND_STD_DEV, _COHERENCE_FACTOR, _TEST_SIGNAL
    int out_dtype;
};

// RxSampleBeamformerTestSignal (RxSampleBeamformerTestSignal.h:25-133).
template <typename T>
T test_signal(bool is3d, T x, T z, T dirx, T diry, T dirz, T d, T dt, T c) {
    T px = ((x + dirx * d) * dt) * c;
    T py = ((diry * d) * dt) * c;
    T pz = is3d ? ((z + dirz * d) * dt) * c : ((dirz * d) * dt) * c;
    T sample = 0;
    for (int k = -3; k <= 3; k++) {
        T ex = px - static_cast<T>(k) * T(6), ey = py - T(30), ez = T(0);
        if (vsqrt(ex * ex + ey * ey + ez * ez) <= T(2)) sample = T(1000);
    }
    if (is3d) {
        for (int k = -3; k <= 3; k++) {
            T ex = T(0), ey = py - T(30), ez = pz - static_cast<T>(k) * T(6);
            if (vsqrt(ex * ex + ey * ey + ez * ez) <= T(2)) sample = T(1000);
        }
    }
    (void)pz;
    return sample;
}

// One pixel of 2D DAS: rxBeamformingDTSPACEKernel (RxBeamformerCuda.cu:230-323)
// with RxSampleBeamformerDelayAndSum::sampleBeamform2D
// (RxSampleBeamformerDelayAndSum.h:122-208) and computeDelayDTSPACE_D
// (RxBeamformerCommon.h:27-33).
// Pinned by: tests/test_oracle_das.py — constant RF (out = k*n), linear ramp
// RF (out = n * weighted mean of the closed-form delays), linearity, a point
// scatterer synthesised in millimetres focusing at its own pixel (pins the
// dt-space scaling and the two-way delay), the TestSignal known answer, FP32
// vs FP64 agreement, brute force on tiny inputs against a pure-Python loop.
template <typename T>
T das_pixel_2d(const DasCall& call, const DtGeometry<T>& G, const T* lut, uint32_t s, uint32_t r) {
    const supra_rx_geometry& g = *call.g;
    const supra_scanline_rx& sc = g.scanlines[s];
    const T dt = static_cast<T>(1.0 / call.fs), c = static_cast<T>(call.c), F = static_cast<T>(call.F);
    const T d = G.depth[r];
    const T aDT = (((d * dt) * c) / (T(2) * F)) / c / dt;
    const T x = G.sx[s], dirx = G.dirx[s], diry = G.diry[s], dirz = G.dirz[s];
    const T inv = T(1) / vmin(aDT, G.medx[s]);
    const T off = static_cast<T>(call.offset);
    const int sel = select_tx(sc);
    T s_interp = 0;
    for (int k = 0; k < 4; k++) {
        bool use = call.interp_tx ? (sc.tx_weights[k] > 0.0) : (k == sel && sc.tx_weights[k] > 0.0);
        if (!use) continue;
        const supra_tx_params& tp = sc.tx[k];
        T s_local;
        if (call.kind == SUPRA_BF_TEST_SIGNAL) {
            s_local = test_signal<T>(false, x, T(0), dirx, diry, dirz, d, dt, c);
        } else {
            const T t0 = G.t0[4 * s + k];
            T sample = 0, wsum = 0;
            int nadd = 0;
            const uint64_t tx_base = static_cast<uint64_t>(tp.tx_scanline_idx) * call.rf.n_ch * call.rf.n_s;
            for (int32_t e = tp.first_active_element[0]; e < tp.last_active_element[0]; e++) {
                uint32_t ch;
                if (g.element_to_channel) {
                    if (g.element_to_channel[e] < 0) continue;
                    ch = static_cast<uint32_t>(g.element_to_channel[e]);
                } else {
                    ch = static_cast<uint32_t>(e) % call.rf.n_ch;
                }
                const T xe = G.ex[e];
                if (std::fabs(xe - x) <= aDT) {
                    const T w = window_get(lut, (xe - x) * inv);
                    wsum += w;
                    nadd++;
                    const T a = xe - (x + dirx * d);
                    const T b = diry * d;
                    T tau = t0 + (vsqrt(a * a + b * b) + d);
                    tau = tau + off;
                    const int32_t i = static_cast<int32_t>(vfloor(tau));
                    const T f = tau - static_cast<T>(i);
                    const uint64_t base = tx_base + static_cast<uint64_t>(ch) * call.rf.n_s;
                    if (static_cast<uint32_t>(i) < call.rf.n_s - 1) {
                        const T v0 = rf_at<T>(call.rf, base + static_cast<uint32_t>(i));
                        const T v1 = rf_at<T>(call.rf, base + static_cast<uint32_t>(i) + 1);
                        sample += w * ((T(1) - f) * v0 + f * v1);
                    } else if (static_cast<uint32_t>(i) < call.rf.n_s && f == T(0)) {
                        sample += w * rf_at<T>(call.rf, base + static_cast<uint32_t>(i));
                    }
                }
            }
            s_local = (nadd > 0) ? (sample / wsum) * static_cast<T>(nadd) : T(0);
            if (call.kind == SUPRA_BF_DELAY_AND_STD_DEV || call.kind == SUPRA_BF_COHERENCE_FACTOR) {
                // RxSampleBeamformerDelayAndStdDev.h:161-270 /
                // RxSampleBeamformerCoherenceFactorDelayAndSum.h:152-258: a second
                // pass over the aperture with the DAS value as `mean` / `resultDas`.
                // An element whose delay is out of range has the value 0 (the
                // reference keeps the previous element's value there).
                const T das_v = s_local;
                T sd = 0, cs = 0, te = 0;
                for (int32_t e = tp.first_active_element[0]; e < tp.last_active_element[0]; e++) {
                    uint32_t ch;
                    if (g.element_to_channel) {
                        if (g.element_to_channel[e] < 0) continue;
                        ch = static_cast<uint32_t>(g.element_to_channel[e]);
                    } else {
                        ch = static_cast<uint32_t>(e) % call.rf.n_ch;
                    }
                    const T xe = G.ex[e];
                    if (!(std::fabs(xe - x) <= aDT)) continue;
                    const T w = window_get(lut, (xe - x) * inv);
                    const T a = xe - (x + dirx * d);
                    const T b = diry * d;
                    T tau = t0 + (vsqrt(a * a + b * b) + d);
                    tau = tau + off;
                    const int32_t i = static_cast<int32_t>(vfloor(tau));
                    const T f = tau - static_cast<T>(i);
                    const uint64_t base = tx_base + static_cast<uint64_t>(ch) * call.rf.n_s;
                    T interp = 0;
                    if (static_cast<uint32_t>(i) < call.rf.n_s - 1)
                        interp = (T(1) - f) * rf_at<T>(call.rf, base + static_cast<uint32_t>(i)) +
                                 f * rf_at<T>(call.rf, base + static_cast<uint32_t>(i) + 1);
                    else if (static_cast<uint32_t>(i) < call.rf.n_s && f == T(0))
                        interp = rf_at<T>(call.rf, base + static_cast<uint32_t>(i));
                    if (call.kind == SUPRA_BF_DELAY_AND_STD_DEV) {
                        const T value = w * interp;
                        sd += (value - das_v) * (value - das_v);
                    } else {
                        cs += interp;
                        te += interp * interp;
                    }
                }
                if (call.kind == SUPRA_BF_DELAY_AND_STD_DEV)
                    s_local = (nadd > 0) ? vsqrt(sd) / wsum * static_cast<T>(nadd) : T(0);
                else
                    s_local = (nadd > 0) ? cs * cs / te / static_cast<T>(nadd) * das_v : T(0);
            }
        }
        if (call.interp_tx) s_interp += static_cast<T>(static_cast<float>(sc.tx_weights[k])) * s_local;
        else s_interp += s_local;
    }
    return clamp_out(s_interp, call.out_dtype);
}

// One pixel of 3D DAS: rxBeamformingDTSPACE3DKernel (RxBeamformerCuda.cu:109-228),
// sampleBeamform3D (RxSampleBeamformerDelayAndSum.h:25-120),
// computeDelayDTSPACE3D_D / computeWindow3D (RxBeamformerCommon.h:35-61).
// Pinned by: the same property tests as 2D run on a matrix-array geometry
// (tests/test_oracle_das.py::test_3d_*), plus the square aperture radius.
template <typename T>
T das_pixel_3d(const DasCall& call, const DtGeometry<T>& G, const T* lut, uint32_t s, uint32_t r) {
    const supra_rx_geometry& g = *call.g;
    const supra_scanline_rx& sc = g.scanlines[s];
    const T dt = static_cast<T>(1.0 / call.fs), c = static_cast<T>(call.c), F = static_cast<T>(call.F);
    const T d = G.depth[r];
    const T t_ap = (((d * dt) * c) / (T(2) * F)) / c / dt;
    const T aDT = t_ap * t_ap;
    const T x = G.sx[s], z = G.sz[s], dirx = G.dirx[s], diry = G.diry[s], dirz = G.dirz[s];
    const T sq = vsqrt(aDT);
    const T invx = T(1) / vmin(sq, G.medx[s]);
    const T invz = T(1) / vmin(sq, G.medz[s]);
    const T off = static_cast<T>(call.offset);
    const int sel = select_tx(sc);
    const uint32_t lx = g.element_layout[0];
    T s_interp = 0;
    for (int k = 0; k < 4; k++) {
        bool use = call.interp_tx ? (sc.tx_weights[k] > 0.0) : (k == sel && sc.tx_weights[k] > 0.0);
        if (!use) continue;
        const supra_tx_params& tp = sc.tx[k];
        T s_local;
        if (call.kind == SUPRA_BF_TEST_SIGNAL) {
            s_local = test_signal<T>(true, x, z, dirx, diry, dirz, d, dt, c);
        } else {
            const T t0 = G.t0[4 * s + k];
            T sample = 0, wsum = 0;
            int nadd = 0;
            const uint64_t tx_base = static_cast<uint64_t>(tp.tx_scanline_idx) * call.rf.n_ch * call.rf.n_s;
            for (uint32_t ex = tp.first_active_element[0]; ex < tp.last_active_element[0]; ex++) {
                for (uint32_t ey = tp.first_active_element[1]; ey < tp.last_active_element[1]; ey++) {
                    const uint32_t e = ex + ey * lx;
                    uint32_t ch;
                    if (g.element_to_channel) {
                        if (g.element_to_channel[e] < 0) continue;
                        ch = static_cast<uint32_t>(g.element_to_channel[e]);
                    } else {
                        ch = e % call.rf.n_ch;
                    }
                    const T xe = G.ex[e], ze = G.ey[e];
                    const T dxs = xe - x, dzs = ze - z;
                    if ((dxs * dxs + dzs * dzs) <= aDT) {
                        const T wxv = window_get(lut, dxs * invx);
                        const T wzv = window_get(lut, dzs * invz);
                        const T w = vsqrt(wxv * wzv);
                        wsum += w;
                        nadd++;
                        const T ax = xe - (x + dirx * d);
                        const T az = ze - (z + dirz * d);
                        const T ay = diry * d;
                        T tau = t0 + (vsqrt((ax * ax + az * az) + ay * ay) + d);
                        tau = tau + off;
                        // static_cast<uint32_t>(floor(.)) on the GPU saturates
                        // negative values to 0 (RxSampleBeamformerDelayAndSum.h:87).
                        const T fl = vfloor(tau);
                        const uint32_t i = (fl > T(0)) ? static_cast<uint32_t>(fl) : 0u;
                        const T f = tau - static_cast<T>(i);
                        const uint64_t base = tx_base + static_cast<uint64_t>(ch) * call.rf.n_s;
                        if (i < call.rf.n_s - 1) {
                            const T v0 = rf_at<T>(call.rf, base + i);
                            const T v1 = rf_at<T>(call.rf, base + i + 1);
                            sample += w * ((T(1) - f) * v0 + f * v1);
                        } else if (i < call.rf.n_s && f == T(0)) {
                            sample += w * rf_at<T>(call.rf, base + i);
                        }
                    }
                }
            }
            s_local = (nadd > 0) ? (sample / wsum) * static_cast<T>(nadd) : T(0);
            if (call.kind == SUPRA_BF_DELAY_AND_STD_DEV || call.kind == SUPRA_BF_COHERENCE_FACTOR) {
                // RxSampleBeamformerDelayAndStdDev.h:36-158 /
                // RxSampleBeamformerCoherenceFactorDelayAndSum.h:33-150 (3D), as in 2D.
                const T das_v = s_local;
                T sd = 0, cs = 0, te = 0;
                for (uint32_t ex = tp.first_active_element[0]; ex < tp.last_active_element[0]; ex++) {
                    for (uint32_t ey = tp.first_active_element[1]; ey < tp.last_active_element[1]; ey++) {
                        const uint32_t e = ex + ey * lx;
                        uint32_t ch;
                        if (g.element_to_channel) {
                            if (g.element_to_channel[e] < 0) continue;
                            ch = static_cast<uint32_t>(g.element_to_channel[e]);
                        } else {
                            ch = e % call.rf.n_ch;
                        }
                        const T xe = G.ex[e], ze = G.ey[e];
                        const T dxs = xe - x, dzs = ze - z;
                        if (!((dxs * dxs + dzs * dzs) <= aDT)) continue;
                        const T w = vsqrt(window_get(lut, dxs * invx) * window_get(lut, dzs * invz));
                        const T ax = xe - (x + dirx * d);
                        const T az = ze - (z + dirz * d);
                        const T ay = diry * d;
                        T tau = t0 + (vsqrt((ax * ax + az * az) + ay * ay) + d);
                        tau = tau + off;
                        const T fl = vfloor(tau);
                        const uint32_t i = (fl > T(0)) ? static_cast<uint32_t>(fl) : 0u;
                        const T f = tau - static_cast<T>(i);
                        const uint64_t base = tx_base + static_cast<uint64_t>(ch) * call.rf.n_s;
                        T interp = 0;
                        if (i < call.rf.n_s - 1)
                            interp = (T(1) - f) * rf_at<T>(call.rf, base + i) + f * rf_at<T>(call.rf, base + i + 1);
                        else if (i < call.rf.n_s && f == T(0))
                            interp = rf_at<T>(call.rf, base + i);
                        if (call.kind == SUPRA_BF_DELAY_AND_STD_DEV) {
                            const T value = w * interp;
                            sd += (value - das_v) * (value - das_v);
                        } else {
                            cs += interp;
                            te += interp * interp;
                        }
                    }
                }
                if (call.kind == SUPRA_BF_DELAY_AND_STD_DEV)
                    s_local = (nadd > 0) ? vsqrt(sd) / wsum * static_cast<T>(nadd) : T(0);
                else
                    s_local = (nadd > 0) ? cs * cs / te / static_cast<T>(nadd) * das_v : T(0);
            }
        }
        if (call.interp_tx) s_interp += static_cast<T>(static_cast<float>(sc.tx_weights[k])) * s_local;
        else s_interp += s_local;
    }
    return clamp_out(s_interp, call.out_dtype);
}

template <typename T>
void das_run(const DasCall& call, const uint32_t* pixels, uint64_t n_pixels, T* out, int nthreads) {
    const supra_rx_geometry& g = *call.g;
    const double dt = 1.0 / call.fs;
    DtGeometry<T> G = to_dt_space<T>(g, dt, call.c);
    T lut[kEntries];
    window_lut<T>(call.window, call.wparam, lut);
    const bool is3d = g.rx_scanline_layout[0] > 1 && g.rx_scanline_layout[1] > 1;
    const uint32_t n_rx = g.num_rx_scanlines;
    const uint64_t total = pixels ? n_pixels : static_cast<uint64_t>(n_rx) * g.num_depths;
    parallel_for(static_cast<int64_t>(total), nthreads, [&](int64_t q) {
        const uint64_t pix = pixels ? pixels[q] : static_cast<uint64_t>(q);
        const uint32_t r = static_cast<uint32_t>(pix / n_rx), s = static_cast<uint32_t>(pix % n_rx);
        out[q] = is3d ? das_pixel_3d<T>(call, G, lut, s, r) : das_pixel_2d<T>(call, G, lut, s, r);
    });
}

// One pixel of the 2D pre-delay: rxDelayDTSPACEKernel + sampleDelay2D
// (src/SupraLib/Beamformer/RawDelay.cu:221-312, 404-468).  Highest-weight
// transmit only, no additional offset; writes w*interp*(n/W) for aperture
// element j = e - first into out[(s*N_ch + j)*N_d + r]; slots j >= last-first
// (never written by the reference, RawDelay.cu:570-572) are defined as 0.
// Pinned by: tests/test_oracle_mv.py::test_rawdelay_sums_to_das — summing the
// channels of the pre-delayed data gives the DAS image (same F-number, window,
// no offset), and a constant RF gives w_e * k * n / W per channel.
template <typename T>
void rawdelay_pixel(const DasCall& call, const DtGeometry<T>& G, const T* lut, uint32_t s, uint32_t r, T* out) {
    const supra_rx_geometry& g = *call.g;
    const supra_scanline_rx& sc = g.scanlines[s];
    const uint32_t n_ch = call.rf.n_ch, n_d = g.num_depths;
    T* dst = out + static_cast<uint64_t>(s) * n_ch * n_d + r;
    for (uint32_t j = 0; j < n_ch; j++) dst[static_cast<uint64_t>(j) * n_d] = 0;
    const T dt = static_cast<T>(1.0 / call.fs), c = static_cast<T>(call.c), F = static_cast<T>(call.F);
    const T d = G.depth[r];
    const T aDT = (((d * dt) * c) / (T(2) * F)) / c / dt;
    const T x = G.sx[s], dirx = G.dirx[s], diry = G.diry[s];
    const T inv = T(1) / vmin(aDT, G.medx[s]);
    const int k = select_tx(sc);
    if (!(sc.tx_weights[k] > 0.0)) return;
    const supra_tx_params& tp = sc.tx[k];
    const T t0 = G.t0[4 * s + k];
    T wsum = 0;
    int nadd = 0;
    for (int32_t e = tp.first_active_element[0]; e < tp.last_active_element[0]; e++) {
        const T xe = G.ex[e];
        if (std::fabs(xe - x) <= aDT) {
            wsum += window_get(lut, (xe - x) * inv);
            nadd++;
        }
    }
    T scale = T(1);
    if (nadd > 0) scale = T(1) / wsum * static_cast<T>(nadd);
    const uint64_t tx_base = static_cast<uint64_t>(tp.tx_scanline_idx) * n_ch * call.rf.n_s;
    uint32_t j = 0;
    for (int32_t e = tp.first_active_element[0]; e < tp.last_active_element[0]; e++, j++) {
        T sample = 0;
        const uint32_t ch = static_cast<uint32_t>(e) % n_ch;
        const T xe = G.ex[e];
        if (std::fabs(xe - x) <= aDT) {
            const T w = window_get(lut, (xe - x) * inv);
            const T a = xe - (x + dirx * d);
            const T b = diry * d;
            const T tau = t0 + (vsqrt(a * a + b * b) + d);
            const int32_t i = static_cast<int32_t>(vfloor(tau));
            const T f = tau - static_cast<T>(i);
            const uint64_t base = tx_base + static_cast<uint64_t>(ch) * call.rf.n_s;
            if (static_cast<uint32_t>(i) < call.rf.n_s - 1) {
                sample = w * ((T(1) - f) * rf_at<T>(call.rf, base + static_cast<uint32_t>(i)) +
                              f * rf_at<T>(call.rf, base + static_cast<uint32_t>(i) + 1));
            } else if (static_cast<uint32_t>(i) < call.rf.n_s && f == T(0)) {
                sample = w * rf_at<T>(call.rf, base + static_cast<uint32_t>(i));
            }
        }
        dst[static_cast<uint64_t>(j) * n_d] = clamp_out(sample * scale, call.out_dtype);
    }
}

// One pixel of the 3D pre-delay: rxDelayDTSPACE3DKernel + sampleDelay3D
// (src/SupraLib/Beamformer/RawDelay.cu:126-220, 315-402).  Highest-weight
// transmit only, no additional offset; aperture radius squared, separable
// window sqrt(w_x w_z) and the saturating floor of the 3D DAS; writes
// w*interp*(n/W) of aperture element (ex, ey) into slot
// (ex - first.x) + (ey - first.y) * element_layout.x of scanline s.  Slots no
// aperture element maps to are 0, and so is an element outside the F-number
// aperture or with its delay out of range: the reference's 3D loop keeps
// `sample` from the previous element there (it is declared outside the
// element loops, RawDelay.cu:150, while the 2D loop resets it per element,
// :264), which is not reproduced.
// Pinned by: tests/test_oracle_mv.py::test_rawdelay3d_sums_to_das3d — the
// channel sum of the pre-delayed volume is the 3D DAS volume (same F-number
// and window, no offset), and a constant RF with a rectangular window gives
// k in every in-aperture slot.
template <typename T>
void rawdelay_pixel_3d(const DasCall& call, const DtGeometry<T>& G, const T* lut, uint32_t s, uint32_t r, T* out) {
    const supra_rx_geometry& g = *call.g;
    const supra_scanline_rx& sc = g.scanlines[s];
    const uint32_t n_ch = call.rf.n_ch, n_d = g.num_depths, lx = g.element_layout[0];
    T* dst = out + static_cast<uint64_t>(s) * n_ch * n_d + r;
    for (uint32_t j = 0; j < n_ch; j++) dst[static_cast<uint64_t>(j) * n_d] = 0;
    const T dt = static_cast<T>(1.0 / call.fs), c = static_cast<T>(call.c), F = static_cast<T>(call.F);
    const T d = G.depth[r];
    const T t_ap = (((d * dt) * c) / (T(2) * F)) / c / dt;
    const T aDT = t_ap * t_ap;
    const T x = G.sx[s], z = G.sz[s], dirx = G.dirx[s], diry = G.diry[s], dirz = G.dirz[s];
    const T sq = vsqrt(aDT);
    const T invx = T(1) / vmin(sq, G.medx[s]);
    const T invz = T(1) / vmin(sq, G.medz[s]);
    const int k = select_tx(sc);
    if (!(sc.tx_weights[k] > 0.0)) return;
    const supra_tx_params& tp = sc.tx[k];
    const T t0 = G.t0[4 * s + k];
    auto weight = [&](uint32_t e, bool& member) {
        const T dxs = G.ex[e] - x, dzs = G.ey[e] - z;
        member = (dxs * dxs + dzs * dzs) <= aDT;
        return member ? vsqrt(window_get(lut, dxs * invx) * window_get(lut, dzs * invz)) : T(0);
    };
    T wsum = 0;
    int nadd = 0;
    for (uint32_t ex = tp.first_active_element[0]; ex < tp.last_active_element[0]; ex++)
        for (uint32_t ey = tp.first_active_element[1]; ey < tp.last_active_element[1]; ey++) {
            bool m;
            const T w = weight(ex + ey * lx, m);
            if (m) {
                wsum += w;
                nadd++;
            }
        }
    T scale = T(1);
    if (nadd > 0) scale = T(1) / wsum * static_cast<T>(nadd);
    const uint64_t tx_base = static_cast<uint64_t>(tp.tx_scanline_idx) * n_ch * call.rf.n_s;
    for (uint32_t ex = tp.first_active_element[0]; ex < tp.last_active_element[0]; ex++)
        for (uint32_t ey = tp.first_active_element[1]; ey < tp.last_active_element[1]; ey++) {
            const uint32_t e = ex + ey * lx;
            const uint32_t slot = (ex - tp.first_active_element[0]) + (ey - tp.first_active_element[1]) * lx;
            if (slot >= n_ch) continue;
            bool m;
            const T w = weight(e, m);
            T sample = 0;
            if (m) {
                const T ax = G.ex[e] - (x + dirx * d);
                const T az = G.ey[e] - (z + dirz * d);
                const T ay = diry * d;
                const T tau = t0 + (vsqrt((ax * ax + az * az) + ay * ay) + d);
                const T fl = vfloor(tau);
                const uint32_t i = (fl > T(0)) ? static_cast<uint32_t>(fl) : 0u;
                const T f = tau - static_cast<T>(i);
                const uint64_t base = tx_base + static_cast<uint64_t>(e % n_ch) * call.rf.n_s;
                if (i < call.rf.n_s - 1)
                    sample = w * ((T(1) - f) * rf_at<T>(call.rf, base + i) + f * rf_at<T>(call.rf, base + i + 1));
                else if (i < call.rf.n_s && f == T(0))
                    sample = w * rf_at<T>(call.rf, base + i);
            }
            dst[static_cast<uint64_t>(slot) * n_d] = clamp_out(sample * scale, call.out_dtype);
        }
}

template <typename T>
void rawdelay_run(const DasCall& call, T* out, int nthreads) {
    const supra_rx_geometry& g = *call.g;
    DtGeometry<T> G = to_dt_space<T>(g, 1.0 / call.fs, call.c);
    T lut[kEntries];
    window_lut<T>(call.window, call.wparam, lut);
    const bool is3d = g.rx_scanline_layout[0] > 1 && g.rx_scanline_layout[1] > 1;
    const uint64_t total = static_cast<uint64_t>(g.num_rx_scanlines) * g.num_depths;
    parallel_for(static_cast<int64_t>(total), nthreads, [&](int64_t q) {
        const uint32_t s = static_cast<uint32_t>(q / g.num_depths), r = static_cast<uint32_t>(q % g.num_depths);
        if (is3d) rawdelay_pixel_3d<T>(call, G, lut, s, r, out);
        else rawdelay_pixel<T>(call, G, lut, s, r, out);
    });
}

// Minimum variance over pre-delayed data x[scanline][channel][sample], FP64,
// with the reference's intended per-sample math over the full depth
// (src/SupraLib/Beamformer/RxBeamformerMV.cu:44-378):
//   aperture = first/last channel with x != 0 (:44-77); L_t = min(L, last-first+1);
//   subarray k active iff first+k+L_t-1 <= last (the offsets of active subarrays
//   are first+k, so the reference's clamp and de-duplication (:98-110) never
//   change them); R(t) = sum_k (x_k - m)(x_k - m)^T over active subarrays
//   (:151-195, means optional :115-149); R~(t) = sum of R(tau) for
//   |tau - t| <= T inside the frame, each R(tau) zero outside its own
//   L_tau x L_tau block, restricted to L_t x L_t (:197-228); loading
//   delta = alpha * trace(R~)/L_t on the diagonal (:239-278; alpha = 1 in the
//   reference); w = R~^-1 1 by Gaussian elimination with partial pivoting
//   (the cuBLAS getrf/getrs of :558-586); y = sum_i (sum_k x[first+k+i]) w_i /
//   ((sum_i w_i) K_act^p); y := 0 if |y| > 1e7 or NaN; out = clampCast(y*N_ch)
//   (:307-378).
// Not reproduced (reference batch bugs for N_s > 2000, :87,146,165,171,173,
// 184-185,361): sample indices and the temporal window are per frame here.
// Pinned by: tests/test_oracle_mv.py — plane-wave closed form
// out = N_ch s(t) K^(1-p) (independent of alpha and T), the alpha -> infinity
// closed form y = sum_i X_i / (L_t K^p), numpy.linalg.solve on the assembled
// systems, and a numpy restatement of R~ on random data.
struct MvCall {
    const void* x;
    int dtype;
    uint32_t n_rx, n_ch, n_d, L, T;
    double p, alpha;
    int means, out_dtype;
    uint32_t pcg_iters = 0;   // 0: LU solve; else the PCG solve (solve_pcg) with at most this many iterations
    double pcg_threshold = 0;
};

inline double x_at(const MvCall& m, uint64_t idx) {
    return m.dtype == SUPRA_INT16 ? static_cast<const int16_t*>(m.x)[idx] : static_cast<const float*>(m.x)[idx];
}

struct MvSample {
    uint32_t first, last, Lt, Kact;
};

// Solve A w = b (n x n, row-major, A and b overwritten) with partial pivoting.
void solve_lu(double* A, double* b, int n) {
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int r = c + 1; r < n; r++)
            if (std::fabs(A[r * n + c]) > std::fabs(A[piv * n + c])) piv = r;
        if (piv != c) {
            for (int k = 0; k < n; k++) std::swap(A[c * n + k], A[piv * n + k]);
            std::swap(b[c], b[piv]);
        }
        const double d = A[c * n + c];
        for (int r = c + 1; r < n; r++) {
            const double f = A[r * n + c] / d;
            for (int k = c; k < n; k++) A[r * n + k] -= f * A[c * n + k];
            b[r] -= f * b[c];
        }
    }
    for (int r = n - 1; r >= 0; r--) {
        double acc = b[r];
        for (int k = r + 1; k < n; k++) acc -= A[r * n + k] * b[k];
        b[r] = acc / A[r * n + r];
    }
}

// RxBeamformerMVpcg's solve (solveBlockwisePCGJacobi, RxBeamformerMVpcg.cu:
// 450-546, reinitialisation frequency 1 as passed at :773-783) in FP64 for
// A w = 1 (n x n, row-major): x0 = 1, r = 1 - A x recomputed every
// iteration, Jacobi preconditioner diag(A), at most max_iter iterations,
// stopping once r.M^-1 r <= 1e-30 or |r| <= threshold.  Returns the
// iterations taken.
// Pinned by: tests/test_oracle_mv.py — with n iterations and no threshold it
// equals the LU solve (conjugate gradients end in n steps), the A-norm of the
// error never grows from one iteration count to the next, and one iteration
// solves a plane-wave system (1 is an eigenvector) exactly.
uint32_t solve_pcg(const double* A, double* w, int n, uint32_t max_iter, double threshold) {
    std::vector<double> x(n, 1.0), r(n), d(n), q(n), sv(n);
    auto matvec = [&](const std::vector<double>& v, std::vector<double>& out) {
        for (int i = 0; i < n; i++) {
            double acc = 0;
            for (int k = 0; k < n; k++) acc += A[i * n + k] * v[k];
            out[i] = acc;
        }
    };
    auto dotp = [&](const std::vector<double>& a, const std::vector<double>& b) {
        double acc = 0;
        for (int i = 0; i < n; i++) acc += a[i] * b[i];
        return acc;
    };
    matvec(x, q);
    for (int i = 0; i < n; i++) r[i] = 1.0 - q[i];
    for (int i = 0; i < n; i++) d[i] = r[i] / A[i * n + i];
    double dnew = dotp(r, d), rnorm = 2 * threshold;
    uint32_t it = 0;
    for (; it < max_iter && dnew > 1e-30 && rnorm > threshold; it++) {
        matvec(d, q);
        const double alpha = dnew / dotp(d, q);
        for (int i = 0; i < n; i++) x[i] += alpha * d[i];
        matvec(x, q);
        for (int i = 0; i < n; i++) r[i] = 1.0 - q[i];
        rnorm = std::sqrt(dotp(r, r));
        for (int i = 0; i < n; i++) sv[i] = r[i] / A[i * n + i];
        const double dold = dnew;
        dnew = dotp(r, sv);
        const double beta = dnew / dold;
        for (int i = 0; i < n; i++) d[i] = sv[i] + beta * d[i];
    }
    for (int i = 0; i < n; i++) w[i] = x[i];
    return it;
}

void mv_scanline(const MvCall& m, uint32_t s, double* out, uint32_t dbg_sample, double* dbg_A, double* dbg_w,
                 int32_t* dbg_meta) {
    const uint32_t n_ch = m.n_ch, n_d = m.n_d, L = m.L;
    const uint32_t K = n_ch - L + 1;
    const uint64_t base = static_cast<uint64_t>(s) * n_ch * n_d;
    auto X = [&](uint32_t ch, uint32_t t) { return x_at(m, base + static_cast<uint64_t>(ch) * n_d + t); };
    std::vector<MvSample> meta(n_d);
    std::vector<double> R(static_cast<size_t>(n_d) * L * L, 0.0);
    std::vector<double> mean(L);
    for (uint32_t t = 0; t < n_d; t++) {
        uint32_t f = 0, l = 0;
        for (uint32_t c = 0; c < n_ch; c++) if (X(c, t) != 0) { f = c; break; }
        for (int32_t c = static_cast<int32_t>(n_ch) - 1; c >= 0; c--) if (X(c, t) != 0) { l = c; break; }
        MvSample ms;
        ms.first = f;
        ms.last = l;
        ms.Lt = std::min(L, l - f + 1);
        ms.Kact = 0;
        for (uint32_t k = 0; k < K; k++) if (f + k + ms.Lt - 1 <= l) ms.Kact++;
        meta[t] = ms;
        std::fill(mean.begin(), mean.end(), 0.0);
        if (m.means) {
            for (uint32_t i = 0; i < ms.Lt; i++) {
                double acc = 0;
                for (uint32_t k = 0; k < ms.Kact; k++) acc += X(f + k + i, t);
                mean[i] = acc / ms.Kact;
            }
        }
        double* Rt = &R[static_cast<size_t>(t) * L * L];
        for (uint32_t k = 0; k < ms.Kact; k++)
            for (uint32_t i = 0; i < ms.Lt; i++) {
                const double xi = X(f + k + i, t) - mean[i];
                for (uint32_t j = 0; j < ms.Lt; j++) Rt[i * L + j] += xi * (X(f + k + j, t) - mean[j]);
            }
    }
    std::vector<double> A(L * L), w(L);
    for (uint32_t t = 0; t < n_d; t++) {
        const MvSample ms = meta[t];
        const uint32_t n = ms.Lt;
        const uint32_t t_lo = t >= m.T ? t - m.T : 0;
        const uint32_t t_hi = std::min(n_d - 1, t + m.T);
        for (uint32_t i = 0; i < n; i++)
            for (uint32_t j = 0; j < n; j++) {
                double acc = 0;
                for (uint32_t tau = t_lo; tau <= t_hi; tau++) acc += R[static_cast<size_t>(tau) * L * L + i * L + j];
                A[i * n + j] = acc;
            }
        double trace = 0;
        for (uint32_t i = 0; i < n; i++) trace += A[i * n + i];
        const double delta = m.alpha * (trace / n);
        for (uint32_t i = 0; i < n; i++) A[i * n + i] += delta;
        if (t == dbg_sample && dbg_A) std::copy(A.begin(), A.begin() + n * n, dbg_A);
        uint32_t iters = 0;
        if (m.pcg_iters) {
            iters = solve_pcg(A.data(), w.data(), static_cast<int>(n), m.pcg_iters, m.pcg_threshold);
        } else {
            for (uint32_t i = 0; i < n; i++) w[i] = 1.0;
            solve_lu(A.data(), w.data(), static_cast<int>(n));
        }
        double wsum = 0;
        for (uint32_t i = 0; i < n; i++) wsum += w[i];
        const double scaling = 1.0 / (wsum * std::pow(static_cast<double>(ms.Kact), m.p));
        double y = 0;
        for (uint32_t i = 0; i < n; i++) {
            double xs = 0;
            for (uint32_t k = 0; k < ms.Kact; k++) xs += X(ms.first + k + i, t);
            y += xs * w[i] * scaling;
        }
        if (std::fabs(y) > 1e7 || std::isnan(y)) y = 0;
        if (t == dbg_sample) {
            if (dbg_w) std::copy(w.begin(), w.begin() + n, dbg_w);
            if (dbg_meta) {
                dbg_meta[0] = static_cast<int32_t>(ms.first);
                dbg_meta[1] = static_cast<int32_t>(ms.last);
                dbg_meta[2] = static_cast<int32_t>(ms.Lt);
                dbg_meta[3] = static_cast<int32_t>(ms.Kact);
                dbg_meta[4] = static_cast<int32_t>(iters);
            }
        }
        out[static_cast<uint64_t>(t) * m.n_rx + s] = clamp_out(y * n_ch, m.out_dtype);
    }
}


// Reverse mode of the 2D DAS (das_pixel_2d) for one frame at speed of sound
// call.c, given dL/dy.  No reference counterpart (the reference does not
// differentiate).  The forward is linear in the RF for fixed c:
//   y = sum_k a_k (n_k/W_k) sum_e w_e [(1-f) rf[i] + f rf[i+1]]   (i+1 < N_s)
// so dL/drf[i] += beta w (1-f), dL/drf[i+1] += beta w f with
// beta = dL/dy * a_k n_k / W_k.  Speed of sound: dt-space lengths scale as
// 1/c while t0 and the offset do not, so dtau/dc = -geo/c with geo the
// geometric part sqrt(q) + d; membership, window and n/W are held fixed.
// Saturated pixels (NaN or beyond FLT_MAX) pass no gradient.
// T is the arithmetic of the values (delays, fractions, sums, n/W); D the
// arithmetic of the discrete decisions (aperture membership, window-table
// index, floor of the delay, range tests, saturation).  T == D is the plain
// FP32 or FP64 restatement; T = double, D = float evaluates the values in
// double precision on the reference's single-precision decisions, which
// separates rounding of the arithmetic from ties that float and double
// resolve differently (e.g. window arguments that are exact multiples of
// 1/63 at pitch multiples).
// norm (optional): a_k n_k / W_k per [slot][scanline][depth], 0 when saturated.
// Pinned by: tests/test_oracle_adjoint.py — the dot-product identity
// <DAS x, y> = <x, DAS^T y> against oracle_das in FP64, dL/dc against FP64
// central differences of oracle_das, brute force of both on a tiny geometry
// against a pure-Python loop; the mixed mode against the FP32 and FP64 modes.
template <typename T>
inline uint32_t window_index_of(T rel) {
    return static_cast<uint32_t>(vround(T(31.5) * (vmin(vmax(rel, T(-1)), T(1)) + T(1))));
}

template <typename T, typename D>
void das_adjoint_frame(const DasCall& call, const double* grad_image, T* grad_rf, double* grad_c, T* norm_out) {
    const supra_rx_geometry& g = *call.g;
    const double dt_d = 1.0 / call.fs;
    DtGeometry<T> G = to_dt_space<T>(g, dt_d, call.c);
    DtGeometry<D> H = to_dt_space<D>(g, dt_d, call.c);
    T lut[kEntries];
    window_lut<T>(call.window, call.wparam, lut);
    const D dt = static_cast<D>(dt_d), c = static_cast<D>(call.c), F = static_cast<D>(call.F);
    const D offD = static_cast<D>(call.offset);
    const T offT = static_cast<T>(call.offset);
    const uint32_t n_rx = g.num_rx_scanlines, n_d = g.num_depths, n_s = call.rf.n_s, n_ch = call.rf.n_ch;
    const int slots = call.interp_tx ? 4 : 1;
    double dc = 0.0;
    // One element-pixel: decisions in D, values in T.
    struct Tap {
        bool member, inr, edge;
        T w, f, geo;
        uint64_t idx;
    };
    auto tap = [&](uint32_t s, int k, int32_t e, uint32_t ch, D d_D, T d_T, D aDT, D inv, Tap& o) {
        const D xeD = H.ex[e], xD = H.sx[s];
        o.member = std::fabs(xeD - xD) <= aDT;
        if (!o.member) return;
        o.w = lut[window_index_of<D>((xeD - xD) * inv)];
        const D aD = xeD - (xD + H.dirx[s] * d_D), bD = H.diry[s] * d_D;
        D tauD = H.t0[4 * s + k] + (vsqrt(aD * aD + bD * bD) + d_D);
        tauD = tauD + offD;
        const int32_t i = static_cast<int32_t>(vfloor(tauD));
        const D fD = tauD - static_cast<D>(i);
        o.inr = static_cast<uint32_t>(i) < n_s - 1;
        o.edge = !o.inr && static_cast<uint32_t>(i) < n_s && fD == D(0);
        const T xeT = G.ex[e], xT = G.sx[s];
        const T aT = xeT - (xT + G.dirx[s] * d_T), bT = G.diry[s] * d_T;
        o.geo = vsqrt(aT * aT + bT * bT) + d_T;
        T tauT = G.t0[4 * s + k] + o.geo;
        tauT = tauT + offT;
        o.f = tauT - static_cast<T>(i);
        const supra_tx_params& tp = g.scanlines[s].tx[k];
        o.idx = (static_cast<uint64_t>(tp.tx_scanline_idx) * n_ch + ch) * n_s + static_cast<uint32_t>(i);
    };
    auto channel = [&](int32_t e, uint32_t& ch) {
        if (g.element_to_channel) {
            if (g.element_to_channel[e] < 0) return false;
            ch = static_cast<uint32_t>(g.element_to_channel[e]);
        } else {
            ch = static_cast<uint32_t>(e) % n_ch;
        }
        return true;
    };
    for (uint32_t s = 0; s < n_rx; s++) {
        const supra_scanline_rx& sc = g.scanlines[s];
        const int sel = select_tx(sc);
        for (uint32_t r = 0; r < n_d; r++) {
            const D d_D = H.depth[r];
            const T d_T = G.depth[r];
            const D aDT = (((d_D * dt) * c) / (D(2) * F)) / c / dt;
            const D inv = D(1) / vmin(aDT, H.medx[s]);
            T norm[4] = {0, 0, 0, 0};
            D s_dec = 0;   // the forward in D, for the saturation decision
            for (int k = 0; k < 4; k++) {
                const bool use = call.interp_tx ? (sc.tx_weights[k] > 0.0) : (k == sel && sc.tx_weights[k] > 0.0);
                if (!use) continue;
                const supra_tx_params& tp = sc.tx[k];
                T wsum = 0;
                D sampleD = 0, wsumD = 0;
                int nadd = 0;
                for (int32_t e = tp.first_active_element[0]; e < tp.last_active_element[0]; e++) {
                    uint32_t ch;
                    if (!channel(e, ch)) continue;
                    Tap t;
                    tap(s, k, e, ch, d_D, d_T, aDT, inv, t);
                    if (!t.member) continue;
                    wsum += t.w;
                    nadd++;
                    const D wD = static_cast<D>(t.w);
                    wsumD += wD;
                    const D fD = static_cast<D>(t.f);
                    if (t.inr)
                        sampleD += wD * ((D(1) - fD) * rf_at<D>(call.rf, t.idx) + fD * rf_at<D>(call.rf, t.idx + 1));
                    else if (t.edge)
                        sampleD += wD * rf_at<D>(call.rf, t.idx);
                }
                const T gk = (nadd > 0) ? static_cast<T>(nadd) / wsum : T(0);
                const T ak = call.interp_tx ? static_cast<T>(static_cast<float>(sc.tx_weights[k])) : T(1);
                norm[k] = call.interp_tx ? ak * gk : gk;
                const D localD = (nadd > 0) ? (sampleD / wsumD) * static_cast<D>(nadd) : D(0);
                s_dec += call.interp_tx ? static_cast<D>(ak) * localD : localD;
            }
            if (!(std::fabs(s_dec) <= D(FLT_MAX)))
                for (int k = 0; k < 4; k++) norm[k] = 0;
            if (norm_out) {
                for (int q = 0; q < slots; q++) {
                    const int k = call.interp_tx ? q : sel;
                    norm_out[(static_cast<uint64_t>(q) * n_rx + s) * n_d + r] = (k >= 0) ? norm[k] : T(0);
                }
            }
            const T gy = static_cast<T>(grad_image[static_cast<uint64_t>(r) * n_rx + s]);
            for (int k = 0; k < 4; k++) {
                if (norm[k] == T(0)) continue;
                const T beta = gy * norm[k];
                const supra_tx_params& tp = sc.tx[k];
                for (int32_t e = tp.first_active_element[0]; e < tp.last_active_element[0]; e++) {
                    uint32_t ch;
                    if (!channel(e, ch)) continue;
                    Tap t;
                    tap(s, k, e, ch, d_D, d_T, aDT, inv, t);
                    if (!t.member) continue;
                    const T bw = beta * t.w;
                    if (t.inr) {
                        if (grad_rf) {
                            grad_rf[t.idx] += bw * (T(1) - t.f);
                            grad_rf[t.idx + 1] += bw * t.f;
                        }
                        const T dv = rf_at<T>(call.rf, t.idx + 1) - rf_at<T>(call.rf, t.idx);
                        dc += static_cast<double>((bw * dv) * t.geo);
                    } else if (t.edge) {
                        if (grad_rf) grad_rf[t.idx] += bw;
                    }
                }
            }
        }
    }
    *grad_c = -dc / call.c;
}

// Pixels of the 2D DAS where two evaluations take different discrete
// decisions for some aperture element (no reference counterpart: a
// classifier for the accuracy analysis of the adjoint).  Evaluation A uses
// speed of sound c_a in precision prec_a (32 or 64), B c_b in prec_b.  The
// decisions are those of das_pixel_2d: aperture membership
// |x_e - x| <= aDT, the window value picked by the rounded table index
// (flagged only where the two indices select different values of the
// single-precision table), the floor of the delay, and the range test on it.
// Flags: 1 membership, 2 window value, 4 floor, 8 range.  Membership and
// window are ratios of dt-space lengths (independent of c in exact
// arithmetic), floor and range are not.
// Pinned by: tests/test_oracle_adjoint.py::test_decision_diff_classifier — a
// rectangular window never sets flag 2, outside the FP32-vs-FP64 flags the
// two images agree to single-precision rounding of the delays and inside them
// the large differences sit, and equal evaluations flag nothing.
struct DecisionEval {
    double c;
    int precision;
};

void das_decision_diff(const DasCall& call, DecisionEval A, DecisionEval B, uint8_t* flags, int nthreads) {
    const supra_rx_geometry& g = *call.g;
    const double dt_d = 1.0 / call.fs;
    const DtGeometry<float> Gfa = to_dt_space<float>(g, dt_d, A.c), Gfb = to_dt_space<float>(g, dt_d, B.c);
    const DtGeometry<double> Gda = to_dt_space<double>(g, dt_d, A.c), Gdb = to_dt_space<double>(g, dt_d, B.c);
    float lutf[kEntries];
    window_lut<float>(call.window, call.wparam, lutf);
    const uint32_t n_rx = g.num_rx_scanlines, n_d = g.num_depths, n_s = call.rf.n_s;
    parallel_for(static_cast<int64_t>(n_rx), nthreads, [&](int64_t si) {
        const uint32_t s = static_cast<uint32_t>(si);
        const supra_scanline_rx& sc = g.scanlines[s];
        const int sel = select_tx(sc);
        for (uint32_t r = 0; r < n_d; r++) {
            uint8_t fl = 0;
            auto eval = [&](auto one, const auto& G, double c_d, int k, int32_t e, bool& member, int& widx,
                            int32_t& fi, bool& inr) {
                using T = decltype(one);
                const T dt = static_cast<T>(dt_d), c = static_cast<T>(c_d), F = static_cast<T>(call.F);
                const T d = G.depth[r];
                const T aDT = (((d * dt) * c) / (T(2) * F)) / c / dt;
                const T inv = T(1) / vmin(aDT, G.medx[s]);
                const T xe = G.ex[e], x = G.sx[s];
                member = std::fabs(xe - x) <= aDT;
                widx = static_cast<int>(window_index_of<T>((xe - x) * inv));
                const T a = xe - (x + G.dirx[s] * d), b = G.diry[s] * d;
                T tau = G.t0[4 * s + k] + (vsqrt(a * a + b * b) + d);
                tau = tau + static_cast<T>(call.offset);
                fi = static_cast<int32_t>(vfloor(tau));
                const T f = tau - static_cast<T>(fi);
                inr = static_cast<uint32_t>(fi) < n_s - 1 || (static_cast<uint32_t>(fi) < n_s && f == T(0));
            };
            auto eval_at = [&](const DecisionEval& E, bool is_a, int k, int32_t e, bool& m, int& w, int32_t& f,
                               bool& in) {
                if (E.precision == 64) eval(0.0, is_a ? Gda : Gdb, E.c, k, e, m, w, f, in);
                else eval(0.0f, is_a ? Gfa : Gfb, E.c, k, e, m, w, f, in);
            };
            for (int k = 0; k < 4; k++) {
                const bool use = call.interp_tx ? (sc.tx_weights[k] > 0.0) : (k == sel && sc.tx_weights[k] > 0.0);
                if (!use) continue;
                const supra_tx_params& tp = sc.tx[k];
                for (int32_t e = tp.first_active_element[0]; e < tp.last_active_element[0]; e++) {
                    if (g.element_to_channel && g.element_to_channel[e] < 0) continue;
                    bool ma, mb, ra, rb;
                    int wa, wb;
                    int32_t fa, fb;
                    eval_at(A, true, k, e, ma, wa, fa, ra);
                    eval_at(B, false, k, e, mb, wb, fb, rb);
                    if (ma != mb) { fl |= 1; continue; }
                    if (!ma) continue;
                    if (wa != wb && lutf[wa] != lutf[wb]) fl |= 2;
                    if (fa != fb) fl |= 4;
                    if (ra != rb) fl |= 8;
                }
            }
            flags[static_cast<uint64_t>(r) * n_rx + s] = fl;
        }
    });
}

// ---- envelope detection and log compression (after beamforming) -----------

// FirFilterFactory::createFilter(length, FilterTypeHilbertTransformer,
// FilterWindowHamming) (src/SupraLib/utilities/FirFilterFactory.h:55-200):
// h[n] = 2/(pi k) for odd k = n - (N-1)/2, else 0, evaluated in double and
// cast, then multiplied by the Hamming window 0.54 - 0.46 cos(2 pi n/(N-1))
// cast to the element type.
// Pinned by: tests/test_oracle_postprocess.py — the closed form in numpy
// float64, antisymmetry h[c-k] = -h[c+k], zero even taps.
template <typename T>
void hilbert_filter(uint32_t n, T* h) {
    const int half = (static_cast<int>(n) - 1) / 2;
    const size_t max_n = n - 1;
    for (uint32_t i = 0; i < n; i++) {
        const int k = static_cast<int>(i) - half;
        h[i] = (k % 2 == 0) ? static_cast<T>(0) : static_cast<T>(2.0 / (M_PI * k));
        h[i] *= static_cast<T>(0.54 - 0.46 * std::cos(2 * M_PI * i / max_n));
    }
}

// kernelFilterDemodulation (src/SupraLib/Beamformer/HilbertFirEnvelope.cu:28-59):
// for sample t of scanline s of an image [sample][scanline], the FIR over
// samples t - N/2 .. t + N/2 inside the image (tap j = t' - (t - N/2)), then
// sqrt(x^2 + acc^2).  Taps past the end of an even-length filter (the
// reference reads one element beyond its filter there) contribute nothing.
// Pinned by: tests/test_oracle_postprocess.py — a pure tone A cos(w t) in
// the filter's passband has the envelope A away from the image ends (the
// analytic signal), a constant image has the envelope |c| (the Hilbert
// transform of a constant is 0 up to the truncated filter's DC leak).
template <typename T>
void envelope_run(const void* in, int in_dtype, uint32_t n_rx, uint32_t n_s, uint32_t n_taps, T* out, int nthreads) {
    std::vector<T> h(n_taps);
    hilbert_filter<T>(n_taps, h.data());
    auto at = [&](uint64_t i) -> T {
        return in_dtype == SUPRA_INT16 ? static_cast<T>(static_cast<const int16_t*>(in)[i])
                                       : static_cast<T>(static_cast<const float*>(in)[i]);
    };
    const int half = static_cast<int>(n_taps) / 2;
    parallel_for(static_cast<int64_t>(n_rx) * n_s, nthreads, [&](int64_t q) {
        const uint32_t t = static_cast<uint32_t>(q / n_rx), s = static_cast<uint32_t>(q % n_rx);
        T acc = 0;
        int j = 0;
        for (int c = static_cast<int>(t) - half; c <= static_cast<int>(t) + half; c++, j++)
            if (c >= 0 && c < static_cast<int>(n_s) && j < static_cast<int>(n_taps))
                acc += at(static_cast<uint64_t>(c) * n_rx + s) * h[j];
        const T v = at(static_cast<uint64_t>(t) * n_rx + s);
        out[q] = vsqrt(v * v + acc * acc);
    });
}

// thrustLogcompress (src/SupraLib/Beamformer/LogCompressor.cu:20-64):
// y = log10(|x| * inScale + 1) * scaleOverDenominator with
// inScale = (float)(10^(dB/20) / inMax) and scaleOverDenominator =
// (float)(gain * outMax / log10(10^(dB/20) + 1)); inMax is first converted
// to the input element type (as the reference's functor argument), outMax is
// 255 for float and uint8 outputs.  T = float evaluates the reference's
// single-precision arithmetic, T = double the intended math.
// Pinned by: tests/test_oracle_postprocess.py — y(0) = 0, y(inMax) = gain *
// 255, and the closed form in numpy float64 on random inputs.
template <typename T>
void logcompress_run(const void* in, int in_dtype, uint64_t count, double dynamic_range_db, double gain,
                     double in_max, T* out, int nthreads) {
    const double dr = std::pow(10.0, dynamic_range_db / 20);
    double in_max_t = in_max;
    if (in_dtype == SUPRA_INT16) in_max_t = static_cast<int16_t>(in_max);
    else if (in_dtype == SUPRA_UINT8) in_max_t = static_cast<uint8_t>(in_max);
    else in_max_t = static_cast<float>(in_max);
    const T in_scale = static_cast<T>(dr / in_max_t);
    const T scale_den = static_cast<T>(gain * 255.0 / std::log10(dr + 1));
    parallel_for(static_cast<int64_t>(count), nthreads, [&](int64_t i) {
        T x;
        if (in_dtype == SUPRA_INT16) x = static_cast<T>(static_cast<const int16_t*>(in)[i]);
        else if (in_dtype == SUPRA_UINT8) x = static_cast<T>(static_cast<const uint8_t*>(in)[i]);
        else x = static_cast<T>(static_cast<const float*>(in)[i]);
        out[i] = std::log10(std::fabs(x) * in_scale + T(1)) * scale_den;
    }, 4096);
}

}  // namespace

extern "C" {

// Window table in FP32 (precision 32) or FP64 (precision 64).
int oracle_window_lut(int window_type, float window_parameter, int precision, void* out) {
    if (!out || window_type < 0 || window_type > 3) return -1;
    if (precision == 64) window_lut<double>(window_type, window_parameter, static_cast<double*>(out));
    else window_lut<float>(window_type, window_parameter, static_cast<float*>(out));
    return 0;
}

// DAS over all pixels ([depth][scanline]) or the listed pixel indices
// (r * num_rx + s).  out: float (precision 32) or double (precision 64).
// sample_beamformer: SUPRA_BF_DELAY_AND_SUM, _DELAY_AND_STD_DEV,
// _COHERENCE_FACTOR or _TEST_SIGNAL.
int oracle_das(const supra_rx_geometry* g, const void* rf, int rf_dtype, uint32_t n_tx, uint32_t n_ch, uint32_t n_s,
               double fs, double f_number, double c_mm_per_s, int window_type, float window_parameter,
               int interpolate_transmits, int32_t additional_offset, int sample_beamformer, int out_dtype, int precision,
               const uint32_t* pixels, uint64_t n_pixels, void* out, int nthreads) {
    if (!g || !rf || !out || n_tx == 0 || n_ch == 0 || n_s == 0) return -1;
    if (sample_beamformer < SUPRA_BF_DELAY_AND_SUM || sample_beamformer > SUPRA_BF_TEST_SIGNAL) return -1;
    for (uint32_t i = 0; i < g->num_rx_scanlines; i++)
        for (int k = 0; k < 4; k++)
            if (g->scanlines[i].tx_weights[k] > 0.0 && g->scanlines[i].tx[k].tx_scanline_idx >= n_tx) return -2;
    DasCall call{g, RfView{rf, rf_dtype, n_tx, n_ch, n_s}, fs, f_number, c_mm_per_s, window_type, window_parameter,
                 interpolate_transmits, additional_offset, sample_beamformer, out_dtype};
    if (precision == 64) das_run<double>(call, pixels, n_pixels, static_cast<double*>(out), nthreads);
    else das_run<float>(call, pixels, n_pixels, static_cast<float*>(out), nthreads);
    return 0;
}


// Pre-delayed RF out[scanline][channel][depth] (2D and 3D geometries), float
// (precision 32) or double (precision 64).
int oracle_rawdelay(const supra_rx_geometry* g, const void* rf, int rf_dtype, uint32_t n_tx, uint32_t n_ch,
                    uint32_t n_s, double fs, double f_number, double c_mm_per_s, int window_type,
                    float window_parameter, int out_dtype, int precision, void* out, int nthreads) {
    if (!g || !rf || !out || n_tx == 0 || n_ch == 0 || n_s == 0) return -1;
    const bool is3d = g->rx_scanline_layout[0] > 1 && g->rx_scanline_layout[1] > 1;
    for (uint32_t i = 0; i < g->num_rx_scanlines; i++) {
        const int k = select_tx(g->scanlines[i]);
        const supra_tx_params& t = g->scanlines[i].tx[k];
        if (g->scanlines[i].tx_weights[k] > 0.0 && t.tx_scanline_idx >= n_tx) return -2;
        if (!is3d && t.last_active_element[0] > t.first_active_element[0] &&
            static_cast<uint32_t>(t.last_active_element[0] - t.first_active_element[0]) > n_ch) return -2;
    }
    DasCall call{g, RfView{rf, rf_dtype, n_tx, n_ch, n_s}, fs, f_number, c_mm_per_s, window_type, window_parameter,
                 0, 0, 0, out_dtype};
    if (precision == 64) rawdelay_run<double>(call, static_cast<double*>(out), nthreads);
    else rawdelay_run<float>(call, static_cast<float*>(out), nthreads);
    return 0;
}

// Minimum variance image out[depth][scanline] (double) from pre-delayed data
// x[scanline][channel][depth].  debug_*: the system (n x n), weights and
// (first, last, L_t, K_act) of sample `debug_sample` of scanline `debug_scanline`.
int oracle_mv(const void* x, int dtype, uint32_t n_rx, uint32_t n_ch, uint32_t n_d, uint32_t sub_array_size,
              uint32_t temporal_smoothing, double scaling_power, int compute_means, double loading_factor,
              int out_dtype, double* out, int nthreads, uint32_t debug_scanline, uint32_t debug_sample,
              double* debug_A, double* debug_w, int32_t* debug_meta) {
    if (!x || !out || n_rx == 0 || n_ch == 0 || n_d == 0) return -1;
    uint32_t L = sub_array_size == 0 ? n_ch / 2 : sub_array_size;
    if (L == 0 || L > n_ch) return -1;
    MvCall m{x, dtype, n_rx, n_ch, n_d, L, temporal_smoothing, scaling_power, loading_factor, compute_means,
             out_dtype};
    parallel_for(static_cast<int64_t>(n_rx), nthreads, [&](int64_t s) {
        const bool dbg = static_cast<uint32_t>(s) == debug_scanline;
        mv_scanline(m, static_cast<uint32_t>(s), out, dbg ? debug_sample : 0xffffffffu, dbg ? debug_A : nullptr,
                    dbg ? debug_w : nullptr, dbg ? debug_meta : nullptr);
    });
    return 0;
}

// The same with RxBeamformerMVpcg's solve (loading trace/L_t; max_iterations
// 0 means the sub-array size, RxBeamformerMVpcg.cu:649).  debug_meta gets a
// fifth entry: the PCG iterations of the debug sample.
int oracle_mvpcg(const void* x, int dtype, uint32_t n_rx, uint32_t n_ch, uint32_t n_d, uint32_t sub_array_size,
                 uint32_t temporal_smoothing, uint32_t max_iterations, double threshold, double scaling_power,
                 int compute_means, int out_dtype, double* out, int nthreads, uint32_t debug_scanline,
                 uint32_t debug_sample, double* debug_A, double* debug_w, int32_t* debug_meta) {
    if (!x || !out || n_rx == 0 || n_ch == 0 || n_d == 0) return -1;
    uint32_t L = sub_array_size == 0 ? n_ch / 2 : sub_array_size;
    if (L == 0 || L > n_ch) return -1;
    MvCall m{x, dtype, n_rx, n_ch, n_d, L, temporal_smoothing, scaling_power, 1.0, compute_means, out_dtype};
    m.pcg_iters = max_iterations ? max_iterations : L;
    // the reference passes the threshold to its kernel as a float (:782): a
    // threshold below the float range is 0, and then no iteration runs
    m.pcg_threshold = static_cast<double>(static_cast<float>(threshold));
    parallel_for(static_cast<int64_t>(n_rx), nthreads, [&](int64_t s) {
        const bool dbg = static_cast<uint32_t>(s) == debug_scanline;
        mv_scanline(m, static_cast<uint32_t>(s), out, dbg ? debug_sample : 0xffffffffu, dbg ? debug_A : nullptr,
                    dbg ? debug_w : nullptr, dbg ? debug_meta : nullptr);
    });
    return 0;
}

// Reverse mode of the 2D DAS for num_frames frames, one speed of sound each
// (host array c_mm_per_s).  grad_image: double [frame][depth][scanline];
// grad_rf: [frame][tx][ch][s] (float or double per precision, overwritten);
// grad_c: double [frame] in 1/(mm/s); norm_out (may be NULL):
// [frame][slot][scanline][depth].  FP32 RF.  Parallel over frames.
// precision: 32 (FP32), 64 (FP64), 6432 (FP64 values on FP32 decisions).
int oracle_das_adjoint(const supra_rx_geometry* g, const float* rf, uint32_t n_frames, uint32_t n_tx, uint32_t n_ch,
                       uint32_t n_s, double fs, double f_number, const double* c_mm_per_s, int window_type,
                       float window_parameter, int interpolate_transmits, int32_t additional_offset,
                       const double* grad_image, int precision, void* grad_rf, double* grad_c, void* norm_out,
                       int nthreads) {
    if (!g || !rf || !c_mm_per_s || !grad_image || !grad_c || n_tx == 0 || n_ch == 0 || n_s == 0) return -1;
    if (g->rx_scanline_layout[0] > 1 && g->rx_scanline_layout[1] > 1) return -3;
    for (uint32_t i = 0; i < g->num_rx_scanlines; i++)
        for (int k = 0; k < 4; k++)
            if (g->scanlines[i].tx_weights[k] > 0.0 && g->scanlines[i].tx[k].tx_scanline_idx >= n_tx) return -2;
    const uint64_t rf_frame = static_cast<uint64_t>(n_tx) * n_ch * n_s;
    const uint64_t img = static_cast<uint64_t>(g->num_rx_scanlines) * g->num_depths;
    const uint64_t norm_frame = img * (interpolate_transmits ? 4 : 1);
    parallel_for(static_cast<int64_t>(n_frames), std::max(1, std::min<int>(nthreads, static_cast<int>(n_frames))),
                 [&](int64_t f) {
        DasCall call{g, RfView{rf + rf_frame * f, SUPRA_FLOAT32, n_tx, n_ch, n_s}, fs, f_number, c_mm_per_s[f],
                     window_type, window_parameter, interpolate_transmits, additional_offset, 0, SUPRA_FLOAT32};
        if (precision == 64 || precision == 6432) {
            double* gr = grad_rf ? static_cast<double*>(grad_rf) + rf_frame * f : nullptr;
            if (gr) std::fill(gr, gr + rf_frame, 0.0);
            double* no = norm_out ? static_cast<double*>(norm_out) + norm_frame * f : nullptr;
            if (precision == 64) das_adjoint_frame<double, double>(call, grad_image + img * f, gr, grad_c + f, no);
            else das_adjoint_frame<double, float>(call, grad_image + img * f, gr, grad_c + f, no);
        } else {
            float* gr = grad_rf ? static_cast<float*>(grad_rf) + rf_frame * f : nullptr;
            if (gr) std::fill(gr, gr + rf_frame, 0.0f);
            das_adjoint_frame<float, float>(call, grad_image + img * f, gr, grad_c + f,
                                            norm_out ? static_cast<float*>(norm_out) + norm_frame * f : nullptr);
        }
    }, 2);
    return 0;
}

// Per-pixel decision differences of the 2D DAS between evaluation A (speed
// of sound c_a, precision prec_a) and B (c_b, prec_b) (das_decision_diff):
// flags uint8 [depth][scanline].
int oracle_das_decision_diff(const supra_rx_geometry* g, uint32_t n_s, double fs, double f_number, double c_a,
                             int prec_a, double c_b, int prec_b, int window_type, float window_parameter,
                             int interpolate_transmits, int32_t additional_offset, uint8_t* flags, int nthreads) {
    if (!g || !flags || n_s == 0) return -1;
    if (g->rx_scanline_layout[0] > 1 && g->rx_scanline_layout[1] > 1) return -3;
    if ((prec_a != 32 && prec_a != 64) || (prec_b != 32 && prec_b != 64)) return -1;
    DasCall call{g, RfView{nullptr, SUPRA_FLOAT32, 0, 0, n_s}, fs, f_number, c_a, window_type,
                 window_parameter, interpolate_transmits, additional_offset, 0, SUPRA_FLOAT32};
    das_decision_diff(call, DecisionEval{c_a, prec_a}, DecisionEval{c_b, prec_b}, flags, nthreads);
    return 0;
}

// Envelope of an image [sample][scanline] (FP32: precision 32, FP64: 64),
// before the output cast.
int oracle_envelope(const void* in, int in_dtype, uint32_t n_rx, uint32_t n_s, uint32_t filter_length,
                    int precision, void* out, int nthreads) {
    if (!in || !out || n_rx == 0 || n_s == 0 || filter_length == 0) return -1;
    if (precision == 64) envelope_run<double>(in, in_dtype, n_rx, n_s, filter_length, static_cast<double*>(out), nthreads);
    else envelope_run<float>(in, in_dtype, n_rx, n_s, filter_length, static_cast<float*>(out), nthreads);
    return 0;
}

// Hilbert FIR coefficients (precision 32 or 64).
int oracle_hilbert_filter(uint32_t filter_length, int precision, void* out) {
    if (!out || filter_length == 0) return -1;
    if (precision == 64) hilbert_filter<double>(filter_length, static_cast<double*>(out));
    else hilbert_filter<float>(filter_length, static_cast<float*>(out));
    return 0;
}

// Log compression of `count` values, before the output cast.
int oracle_logcompress(const void* in, int in_dtype, uint64_t count, double dynamic_range_db, double gain,
                       double in_max, int precision, void* out, int nthreads) {
    if (!in || !out) return -1;
    if (precision == 64)
        logcompress_run<double>(in, in_dtype, count, dynamic_range_db, gain, in_max, static_cast<double*>(out), nthreads);
    else
        logcompress_run<float>(in, in_dtype, count, dynamic_range_db, gain, in_max, static_cast<float*>(out), nthreads);
    return 0;
}

}  // extern "C"
