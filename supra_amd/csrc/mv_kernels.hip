// Minimum-variance receive beamforming over pre-delayed data, fused into one
// kernel for gfx950, and the supra_mv_run entry point.
//
// Stands in for RxBeamformerMV::performRxBeamforming
// (src/SupraLib/Beamformer/RxBeamformerMV.cu:382-614) and its kernels
// (:79-378) plus the cuBLAS batched LU solve (:558-586).  The reference walks
// every scanline and 2000-sample batch from the host (≥9 launches and 3
// device-wide syncs per batch); here one launch covers the frame:
//
//   workgroup = (scanline, tile of Dt output depths), 256 lanes
//   1. stage x[channel][tau] for the tile and its +-T halo in LDS (float),
//      read or, in the fused mode, computed from the RF; find each sample's
//      aperture (first/last non-zero channel, :44-77, one wave and 64-channel
//      ballots per sample), the subarray length L_t and the number of active
//      subarrays (:79-113);
//   2. R(tau) = sum_k (x_k - m)(x_k - m)^T for every halo sample, upper
//      triangle packed in LDS, by diagonals: a K-term FP64 sum for the first
//      entry of each diagonal, then the FP64 recurrence that adds the entering
//      and removes the leaving product (:151-195); each lane walks a diagonal
//      pair (d, L-1-d) so the trip counts match;
//   3. one wave (L <= 64) or half-wave (L <= 32) per output sample: row i of
//      R~ = sum over |tau - t| <= T (:197-228) in lane i's registers,
//      diagonal loading alpha*trace/L_t (:239-305), a Cholesky factorisation
//      with FP64 pivots and diagonal, triangular solves for w = R~^-1 1, and,
//      for samples that may be ill-conditioned (condition bound 1 + L_t/alpha
//      of the loaded system and squared-pivot ratio both large), an FP64
//      residual check with up to two steps of iterative refinement (FP32
//      factors, FP64 residual); then the output
//      y = sum_i (sum_k x[f+k+i]) w_i / (sum w * K^p) (:307-378).
//
// Numerics are FP32 with FP64 only where stated; results are compared with
// the FP64 host oracle (per-sample relative error), not bit for bit: the
// reference's own sums are atomics in arbitrary order and its solve is
// cuBLAS, so its bits are not reproducible either.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>

#include "bf_internal.hpp"
#include "das_device.hpp"
#include "mv_launch.hpp"
#include "status.hpp"
#include "supra_bf.h"

namespace supra_amd {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxTile = 64;


template <typename T> __device__ inline float in_f(T v) { return static_cast<float>(v); }

template <typename T> __device__ inline T clamp_out(float x);
template <> __device__ inline float clamp_out<float>(float x) { return fminf(fmaxf(x, -FLT_MAX), FLT_MAX); }
template <> __device__ inline int16_t clamp_out<int16_t>(float x) {
    return static_cast<int16_t>(fminf(fmaxf(x, -32767.0f), 32767.0f));
}

__device__ inline float wave_sum(float v) {
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}
__device__ inline double wave_sum(double v) {
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline float bcast(float v, int lane) { return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane)); }
__device__ inline double bcast(double v, int lane) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(u), lane));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(u >> 32), lane));
    return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

struct SampleMeta {
    int32_t first, lt, kact;
    int32_t bad;    // a non-finite channel value: every window holding it ends as NaN -> 0
};

__device__ inline uint32_t packed_index(uint32_t i, uint32_t j, uint32_t L) {
    // upper triangle, row-major: row i holds columns i..L-1
    return i * L - (i * (i - 1)) / 2 + (j - i);
}

// Sub-wave helpers: with PPW == 2 pixels per wave, lanes [0, 32) and
// [32, 64) each hold one 32 x 32 system (row i in lane i of the half).
template <int PPW>
__device__ inline float sub_bcast(float v, int j, int half) {
    if (PPW == 1) return bcast(v, j);
    const float a0 = bcast(v, j), a1 = bcast(v, j + 32);
    return half ? a1 : a0;
}
template <int PPW>
__device__ inline double sub_bcast(double v, int j, int half) {
    if (PPW == 1) return bcast(v, j);
    const double a0 = bcast(v, j), a1 = bcast(v, j + 32);
    return half ? a1 : a0;
}
// Broadcast of lane j (0..31) within each half-wave by ds_swizzle (bit-mask
// mode: and 0, or j), one instruction instead of two readlanes and a select;
// j must fold to a constant (unrolled loops).
template <int J> __device__ inline int swizzle_from(int v) { return __builtin_amdgcn_ds_swizzle(v, J << 5); }
__device__ inline float half_bcast(float v, int j) {
    const int x = __float_as_int(v);
    int r = 0;
    switch (j) {
        case 0: r = swizzle_from<0>(x); break;
        case 1: r = swizzle_from<1>(x); break;
        case 2: r = swizzle_from<2>(x); break;
        case 3: r = swizzle_from<3>(x); break;
        case 4: r = swizzle_from<4>(x); break;
        case 5: r = swizzle_from<5>(x); break;
        case 6: r = swizzle_from<6>(x); break;
        case 7: r = swizzle_from<7>(x); break;
        case 8: r = swizzle_from<8>(x); break;
        case 9: r = swizzle_from<9>(x); break;
        case 10: r = swizzle_from<10>(x); break;
        case 11: r = swizzle_from<11>(x); break;
        case 12: r = swizzle_from<12>(x); break;
        case 13: r = swizzle_from<13>(x); break;
        case 14: r = swizzle_from<14>(x); break;
        case 15: r = swizzle_from<15>(x); break;
        case 16: r = swizzle_from<16>(x); break;
        case 17: r = swizzle_from<17>(x); break;
        case 18: r = swizzle_from<18>(x); break;
        case 19: r = swizzle_from<19>(x); break;
        case 20: r = swizzle_from<20>(x); break;
        case 21: r = swizzle_from<21>(x); break;
        case 22: r = swizzle_from<22>(x); break;
        case 23: r = swizzle_from<23>(x); break;
        case 24: r = swizzle_from<24>(x); break;
        case 25: r = swizzle_from<25>(x); break;
        case 26: r = swizzle_from<26>(x); break;
        case 27: r = swizzle_from<27>(x); break;
        case 28: r = swizzle_from<28>(x); break;
        case 29: r = swizzle_from<29>(x); break;
        case 30: r = swizzle_from<30>(x); break;
        case 31: r = swizzle_from<31>(x); break;
    }
    return __int_as_float(r);
}
// Broadcast for the throughput-bound loops (solves, residual): ds_swizzle
// with two systems per wave, readlane with one.
template <int PPW>
__device__ inline float sub_bcast_swz(float v, int j, int half) {
    if (PPW == 1) return bcast(v, j);
    (void)half;
    return half_bcast(v, j);
}
template <int PPW>
__device__ inline float sub_sum(float v) {
    for (int m = 32 / PPW; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}
template <int PPW>
__device__ inline double sub_sum(double v) {
    for (int m = 32 / PPW; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}
template <int PPW>
__device__ inline float sub_max(float v) {
    for (int m = 32 / PPW; m > 0; m >>= 1) v = fmaxf(v, __shfl_xor(v, m, 64));
    return v;
}
template <int PPW>
__device__ inline float sub_min(float v) {
    for (int m = 32 / PPW; m > 0; m >>= 1) v = fminf(v, __shfl_xor(v, m, 64));
    return v;
}

// Conditioning below which the FP64 residual check is skipped: the a-priori
// bound 1 + n/alpha of a loaded system, or the squared-pivot ratio.
constexpr float kKappaSafe = 1.0e4f;
constexpr float kPivotRatioSafe = 64.0f;

// Solve (G G^T) v = b for the sub-wave's pixel: G's row i is in lane i's
// registers a[] (strictly lower part), ginv is lane i's 1/G[i][i]; the
// strictly lower factor is also in LDS (row-major packed, G[r][c] at
// r (r - 1) / 2 + c) for the transposed solve.  Lanes i >= n carry no data.
template <int LMAX, int PPW>
__device__ inline float chol_solve(const float (&a)[LMAX], float ginv, float b, int n, int li, int half,
                                   const float* lmat) {
    // fresh lane-vs-step predicates for each call (not kept live across calls)
    asm volatile("" : "+v"(li), "+v"(n));
    // forward: G z = b, column-oriented
    float c = b, z = 0.0f;
#pragma unroll
    for (int j = 0; j < LMAX; j++) {
        const float zj = sub_bcast_swz<PPW>(c * ginv, j, half);
        if (li == j) z = zj;
        if (li > j && li < n && j < n) c = fmaf(-a[j], zj, c);
    }
    // backward: G^T v = z, column j of G^T is row j of G (in LDS)
    float v = 0.0f;
    c = z;
#pragma unroll
    for (int j = LMAX - 1; j >= 0; j--) {
        const float vj = sub_bcast_swz<PPW>(c * ginv, j, half);
        if (li == j) v = vj;
        if (li < j && j < n) c = fmaf(-lmat[(j * (j - 1)) / 2 + li], vj, c);
    }
    return v;
}

// Jacobi-preconditioned conjugate gradients for R~ w = 1 on the sub-wave's
// pixel, as solveBlockwisePCGJacobi (RxBeamformerMVpcg.cu:450-546) with the
// reinitialisation frequency of 1 that performRxBeamforming passes (:773-783):
// x0 = 1, the residual recomputed as 1 - R~ x every iteration, at most
// max_iter iterations, stopping once r.M^-1 r <= 1e-30 or |r| <= threshold
// (M = diag R~).  Lane i holds row i of R~ in A[] (entries k >= n unused)
// and its diagonal in mdiag; lanes i >= n carry no data.  Both sub-waves of
// a wave run the same number of steps; a finished one keeps its x.
template <int LMAX, int PPW>
__device__ inline float pcg_solve(const float (&A)[LMAX], float mdiag, int n, int li, int half, uint32_t max_iter,
                                  float threshold) {
    const bool mine = li < n;
    auto matvec = [&](float v) {
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < LMAX; k++) {
            const float vk = sub_bcast_swz<PPW>(v, k, half);
            if (k < n) acc = fmaf(A[k], vk, acc);
        }
        return acc;
    };
    float x = mine ? 1.0f : 0.0f;
    const float ax0 = matvec(x);
    float r = mine ? 1.0f - ax0 : 0.0f;
    float d = mine ? r / mdiag : 0.0f;
    float dnew = sub_sum<PPW>(mine ? r * d : 0.0f);
    float rnorm = 2.0f * threshold;
    bool active = max_iter > 0 && dnew > 1e-30f && rnorm > threshold;
    for (uint32_t it = 0; __any(active);) {
        const float q = matvec(d);
        const float alpha = dnew / sub_sum<PPW>(mine ? d * q : 0.0f);
        if (active) x = fmaf(alpha, d, x);
        const float ax = matvec(x);
        if (active) r = mine ? 1.0f - ax : 0.0f;
        const float rr = sub_sum<PPW>(mine ? r * r : 0.0f);
        const float sv = mine ? r / mdiag : 0.0f;
        const float dn = sub_sum<PPW>(mine ? r * sv : 0.0f);
        if (active) {
            rnorm = sqrtf(rr);
            const float beta = dn / dnew;
            dnew = dn;
            d = fmaf(beta, d, sv);
        }
        it++;
        active = active && it < max_iter && dnew > 1e-30f && rnorm > threshold;
    }
    return x;
}

// Fused mode: the pre-delayed samples of scanline s, depths [h_lo, h_lo+H),
// computed from the RF exactly as the pre-delay kernel writes them
// (rawdelay2d_kernel in das_kernels.hip; RawDelay.cu:221-312, 404-468):
// highest-weight transmit, aperture element c = e - first, w * interp * n/W
// through clampCast<float>, 0 for c >= last - first.  Same operations in the
// same order, so the values are bit-identical to the two-kernel path.
template <typename RFT>
__device__ inline void stage_delayed(const MvArgs& a, const float* lut, uint32_t s, uint32_t h_lo, int H,
                                     float* xs, unsigned char* lds) {
    const DasArgs& g = a.g;
    const int n_ch = static_cast<int>(a.n_ch);
    const int hstride = static_cast<int>(a.tile + 2 * a.T);
    float* s_lut = reinterpret_cast<float*>(lds + a.off_fused);
    float* s_xe = s_lut + kWindowEntries;
    float* s_dx = s_xe + n_ch;
    float* s_dx2 = s_dx + n_ch;
    uint32_t* s_off = reinterpret_cast<uint32_t*>(s_dx2 + n_ch);
    float* p_d = reinterpret_cast<float*>(s_off + n_ch);
    float* p_adt = p_d + hstride;
    float* p_inv = p_adt + hstride;
    float* p_px = p_inv + hstride;
    float* p_py2 = p_px + hstride;
    float* p_scale = p_py2 + hstride;

    const ScanView sc = scan_view(g.scans, s);
    const int k = sc.sel();
    TxDT tx{};
    int cnt = 0;
    if (k >= 0) {
        tx = sc.tx(k);
        cnt = tx.last_x > tx.first_x ? static_cast<int>(tx.last_x - tx.first_x) : 0;
    }
    for (int i = threadIdx.x; i < kWindowEntries; i += kThreads) s_lut[i] = lut[i];
    for (int j = threadIdx.x; j < cnt; j += kThreads) {
        const int e = static_cast<int>(tx.first_x) + j;
        const float xe = g.ex[e];
        const float dx = xe - sc.x();
        s_xe[j] = xe;
        s_dx[j] = dx;
        s_dx2[j] = dx * dx;
        s_off[j] = (static_cast<uint32_t>(e) % g.n_ch) * g.n_s;
    }
    __syncthreads();
    // per-sample geometry and the aperture's weight sum and count
    for (int h = threadIdx.x; h < H; h += kThreads) {
        const float d = g.depths[h_lo + h];
        const float aDT = (((d * g.dt) * g.c) / (2.0f * g.F)) / g.c / g.dt;
        const float inv = 1.0f / fminf(aDT, sc.max_ed_x());
        const float py = sc.dir_y() * d;
        float wsum = 0.0f;
        int nadd = 0;
        for (int j = 0; j < cnt; j++) {
            const float dx = s_dx[j];
            if (fabsf(dx) <= aDT) {
                wsum += s_lut[window_index(dx * inv)];
                nadd++;
            }
        }
        float scale = 1.0f;
        if (nadd > 0) scale = 1.0f / wsum * static_cast<float>(nadd);
        p_d[h] = d;
        p_adt[h] = aDT;
        p_inv[h] = inv;
        p_px[h] = sc.x() + sc.dir_x() * d;
        p_py2[h] = py * py;
        p_scale[h] = scale;
    }
    __syncthreads();
    const bool vertical = (sc.dir_x() == 0.0f);
    const RFT* rf_tx = static_cast<const RFT*>(g.rf) + static_cast<uint64_t>(tx.tx_idx) * g.n_ch * g.n_s;
    const __amdgpu_buffer_rsrc_t rsrc = rf_resource(rf_tx, g.n_ch * g.n_s * sizeof(RFT));
    const int rs = static_cast<int>(a.rs);
#pragma unroll 4
    for (int idx = threadIdx.x; idx < H * rs; idx += kThreads) {
        const int c = idx / H, h = idx % H;
        float v = 0.0f;
        if (c < cnt) {
            float sample = 0.0f;
            const float dx = s_dx[c];
            if (fabsf(dx) <= p_adt[h]) {
                const float w = s_lut[window_index(dx * p_inv[h])];
                float q;
                if (vertical) {
                    q = s_dx2[c] + p_py2[h];
                } else {
                    const float e = s_xe[c] - p_px[h];
                    q = e * e + p_py2[h];
                }
                const float tau = tx.t0 + (sqrt_rn(q) + p_d[h]);
                gather<RFT, false>(rsrc, s_off[c], g.n_s, tau, w, sample);
            }
            v = clamp_cast<float>(sample * p_scale[h]);
        }
        xs[h * rs + c] = v;
    }
    __syncthreads();
}

// One workgroup tile: output depths [tile * a.tile, ...) of scanline s.
// Diagonals da <= db of one sample's R (db == da: one diagonal), entries
// (i, i + d) for i < L - d, given each first entry's sum
// r = sum_{k < K_act} x[f+k] x[f+k+d] (FP64): the rest by the recurrence,
// minus K_act m_i m_{i+d} with means, single-precision stores; entries beyond
// the sample's L_t are 0.  One loop walks diagonal da then diagonal db, so
// the trip count (L_t + 1 entries for a pair) is the same on every lane that
// holds a full-length sample.  Returns whether a stored entry is not finite.
template <bool MEANS>
__device__ inline bool cov_diagonals(const float* row, const float* mean, float* dst, int L, int lt, int kact, int da,
                                     int db, double ra, double rb) {
    bool nonfinite = false;
    const bool two = db != da;
    const int na = max(lt - da, 0), nb = two ? max(lt - db, 0) : 0;
    double r = ra;
    int d = da, i = 0;
    for (int step = 0; step < na + nb; step++, i++) {
        if (step == na) {   // on to the second diagonal
            r = rb;
            d = db;
            i = 0;
        }
        if (i > 0) {
            r = fma(-static_cast<double>(row[i - 1]), static_cast<double>(row[i - 1 + d]), r);
            r = fma(static_cast<double>(row[kact - 1 + i]), static_cast<double>(row[kact - 1 + i + d]), r);
        }
        double v = r;
        if (MEANS) v -= static_cast<double>(kact) * static_cast<double>(mean[i]) * static_cast<double>(mean[i + d]);
        const float fv = static_cast<float>(v);
        nonfinite |= !(fabsf(fv) <= FLT_MAX);
        dst[packed_index(i, i + d, L)] = fv;
    }
    for (int j = na; j + da < L; j++) dst[packed_index(j, j + da, L)] = 0.0f;
    if (two)
        for (int j = nb; j + db < L; j++) dst[packed_index(j, j + db, L)] = 0.0f;
    return nonfinite;
}

template <typename InT, typename OutT, int LMAX, bool MEANS, bool FUSED>
__device__ __attribute__((always_inline)) inline void mv_tile(const MvArgs& a, const float* lut, uint32_t s, uint32_t tile, unsigned char* lds) {
    float* xs = reinterpret_cast<float*>(lds);
    SampleMeta* meta = reinterpret_cast<SampleMeta*>(lds + a.off_meta);
    float* mean = reinterpret_cast<float*>(lds + a.off_mean);
    float* rp = reinterpret_cast<float*>(lds + a.off_r);

    const uint32_t t0 = tile * a.tile;
    const uint32_t t1 = min(t0 + a.tile, a.n_d);          // output depths [t0, t1)
    const uint32_t h_lo = t0 >= a.T ? t0 - a.T : 0;
    const uint32_t h_hi = min(a.n_d - 1, t1 - 1 + a.T);   // halo samples [h_lo, h_hi]
    const int H = static_cast<int>(h_hi - h_lo + 1);
    const int L = static_cast<int>(a.L);
    const int rs = static_cast<int>(a.rs);

    // 1. stage x (zero padded to rs channels) and per-sample aperture
    if (FUSED) {
        stage_delayed<InT>(a, lut, s, h_lo, H, xs, lds);
    } else {
        const InT* xin = static_cast<const InT*>(a.x) + static_cast<uint64_t>(s) * a.n_ch * a.n_d;
        // batches of 8 loads per lane in flight before the LDS stores (the
        // loop is otherwise one global-memory round trip per element)
        constexpr int kBatch = 8;
        const int total = H * rs, n_ch = static_cast<int>(a.n_ch);
        for (int base = 0; base < total; base += kBatch * kThreads) {
            float v[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; u++) {
                const int idx = base + u * kThreads + static_cast<int>(threadIdx.x);
                const int c = idx / H, h = idx % H;
                v[u] = (idx < total && c < n_ch) ? in_f(xin[static_cast<uint64_t>(c) * a.n_d + h_lo + h]) : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < kBatch; u++) {
                const int idx = base + u * kThreads + static_cast<int>(threadIdx.x);
                const int c = idx / H, h = idx % H;
                if (idx < total) xs[h * rs + c] = v[u];
            }
        }
        __syncthreads();
    }
    // first / last channel with a non-zero sample (0 when there is none): one
    // wave per sample, 64 channels per ballot
    {
        const int wv = threadIdx.x / 64, ln = threadIdx.x % 64;
        const int n_ch = static_cast<int>(a.n_ch);
        for (int h = wv; h < H; h += kWaves) {
            const float* row = xs + h * rs;
            int f = -1, l = 0;
            for (int cb = 0; cb < n_ch; cb += 64) {
                const int c = cb + ln;
                const uint64_t nz = __ballot(c < n_ch && row[c] != 0.0f);
                if (nz != 0) {
                    if (f < 0) f = cb + __ffsll(static_cast<unsigned long long>(nz)) - 1;
                    l = cb + 63 - __clzll(static_cast<long long>(nz));
                }
            }
            if (f < 0) f = 0;
            if (ln == 0) {
                SampleMeta m;
                m.first = f;
                m.lt = min(L, l - f + 1);
                m.kact = min(static_cast<int>(a.K), l - f - m.lt + 2);
                m.bad = 0;
                meta[h] = m;
            }
        }
    }
    __syncthreads();
    if (MEANS) {
        for (int idx = threadIdx.x; idx < H * L; idx += kThreads) {
            const int h = idx / L, i = idx % L;
            const SampleMeta m = meta[h];
            float acc = 0.0f;
            if (i < m.lt) {
                for (int k = 0; k < m.kact; k++) acc += xs[h * rs + m.first + k + i];
                acc = acc / static_cast<float>(m.kact);
            }
            mean[h * L + i] = acc;
        }
        __syncthreads();
    }

    // 2. R(tau) per halo sample by diagonals: with x_k = x[f+k .. f+k+L_t),
    //    R[i][j] = sum_{k < K_act} x[f+k+i] x[f+k+j] satisfies
    //    R[i][j] = R[i-1][j-1] - x[f+i-1] x[f+j-1] + x[f+K_act-1+i] x[f+K_act-1+j],
    //    so each diagonal d = j - i costs one K_act-term sum and two products
    //    per further entry: O(L K + L^2) instead of O(L^2 K).  Products of two
    //    floats are exact in double; the sums and the recurrence run in FP64,
    //    so the stored single-precision entries carry one rounding each.
    //    Means: sum (x_k - m)(x_k - m)^T = sum x_k x_k^T - K_act m m^T.
    //    Each lane takes the diagonal pair (d, L-1-d) of one sample, so every
    //    lane walks about the same number of entries, and the two first-entry
    //    sums share their loads of x[f+k] (two independent FMA chains).
    const int pairs = (L + 1) / 2;
    for (int item = threadIdx.x; item < H * pairs; item += kThreads) {
        const int h = item / pairs, da = item % pairs, db = L - 1 - da;   // db == da: the middle diagonal
        const SampleMeta m = meta[h];
        const float* row = xs + h * rs + m.first;
        float* dst = rp + h * a.packed;
        const int lt = m.lt, kact = m.kact;
        double ra = 0.0, rb = 0.0;
        if (db < lt) {
            for (int k = 0; k < kact; k++) {
                const double xk = static_cast<double>(row[k]);
                ra = fma(xk, static_cast<double>(row[k + da]), ra);
                rb = fma(xk, static_cast<double>(row[k + db]), rb);
            }
        } else if (da < lt) {
            for (int k = 0; k < kact; k++) ra = fma(static_cast<double>(row[k]), static_cast<double>(row[k + da]), ra);
        }
        const bool nonfinite = cov_diagonals<MEANS>(row, mean + h * L, dst, L, lt, kact, da, db, ra, rb);
        // An overflowed or NaN covariance (e.g. from a -FLT_MAX pre-delayed
        // sample) makes every window holding the sample NaN in single
        // precision, hence output 0; it is kept out of the window sums.
        if (nonfinite) meta[h].bad = 1;
    }
    __syncthreads();

    const int packed = static_cast<int>(a.packed);

    // 3. per output sample: one wave (PPW == 1) or one half-wave (PPW == 2)
    constexpr int PPW = (LMAX <= 32) ? 2 : 1;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (wave >= static_cast<int>(a.waves)) return;   // (no barrier follows in this tile)
    const int half = (PPW == 2) ? lane / 32 : 0;
    const int li_lane = (PPW == 2) ? lane % 32 : lane;
    // strictly lower factor, row-major packed: G[r][c] (c < r) at r (r - 1) / 2 + c
    float* lmat = reinterpret_cast<float*>(lds + a.off_wave + (wave * PPW + half) * a.wave_bytes);
    const int lrow_lane = (li_lane * (li_lane - 1)) / 2;
    OutT* out = static_cast<OutT*>(a.out);
    for (uint32_t tb = t0 + wave * PPW; tb < t1; tb += a.waves * PPW) {
        // The lane index is made opaque per sample: otherwise the compiler
        // hoists every lane-vs-step comparison of the unrolled factorisation
        // and solves out of this loop as a 64-bit mask each, and spills them.
        int li = li_lane, lrow = lrow_lane;
        asm volatile("" : "+v"(li), "+v"(lrow));
        const bool valid = tb + half < t1;
        const uint32_t t = valid ? tb + half : t1 - 1;    // a spare half repeats the last sample
        const int ht = static_cast<int>(t - h_lo);
        const SampleMeta m = meta[ht];
        const int n = m.lt;
        const int w_lo = static_cast<int>((t >= a.T ? t - a.T : 0) - h_lo);
        const int w_hi = static_cast<int>(min(a.n_d - 1, t + a.T) - h_lo);
        // row `li` of R~ = sum of R(tau) over the window (upper-packed source
        // mirrored); the window loop is outside so that each step issues the
        // row's LDS loads back to back.  A window holding a non-finite
        // covariance (see phase 2) ends as NaN in single precision: output 0.
        // Entries k >= n of a row, and the rows of lanes li >= n, are never
        // read by the solve below, so the sums run unmasked over clamped
        // in-range indices: one LDS load and one add per entry, no per-lane
        // predicates to keep live across the window loop.
        float row[LMAX];
#pragma unroll
        for (int k = 0; k < LMAX; k++) row[k] = 0.0f;
        const int lic = min(li, L - 1);
        int pidx[LMAX];
#pragma unroll
        for (int k = 0; k < LMAX; k++) {
            const int kc = min(k, L - 1);
            const int i0 = min(lic, kc), j0 = max(lic, kc);
            pidx[k] = i0 * L - (i0 * (i0 - 1)) / 2 + (j0 - i0);
        }
        // The diagonal entry is summed a second time on its own (same terms,
        // same order, so diag == row[li] for li < n): extracting row[li]
        // with a lane-dependent index would demote the row to scratch memory.
        const int pdiag = lic * L - (lic * (lic - 1)) / 2;
        float diag = 0.0f;
        int bad = 0;
        for (int h = w_lo; h <= w_hi; h++) {
            const SampleMeta mh = meta[h];
            bad |= mh.bad;
            if (mh.bad) continue;
            const float* src = rp + h * packed;
#pragma unroll
            for (int k = 0; k < LMAX; k++) row[k] += src[pidx[k]];
            diag += src[pdiag];
        }
        const float trace = sub_sum<PPW>(li < n ? diag : 0.0f);
        const float delta = ((1.0f / static_cast<float>(n)) * trace) * a.alpha;
        // R~ row li with the loading on its diagonal (adding 0.0f elsewhere
        // leaves the entries unchanged)
        float orig[LMAX];
#pragma unroll
        for (int k = 0; k < LMAX; k++) orig[k] = row[k] + ((k == li) ? delta : 0.0f);
        // w = R~^-1 1: Cholesky (RxBeamformerMV) or, with pcg_iters > 0,
        // Jacobi-preconditioned conjugate gradients (RxBeamformerMVpcg)
        float w;
        if (a.pcg_iters == 0) {
            // Cholesky, row `li` in registers, FP64 pivots and diagonal.  Row
            // entries k >= li are never read by the factorisation, so the loading
            // enters only through the diagonal dd.
            double dd = static_cast<double>(diag + delta);
            // steps beyond the larger system of the wave are skipped (uniform branch)
            const int nmax = (PPW == 1) ? n : max(__builtin_amdgcn_readlane(n, 0), __builtin_amdgcn_readlane(n, 32));
    #pragma unroll
            for (int j = 0; j < LMAX; j++) {
                if (j < nmax) {   // a guard, not a break: the loop must unroll fully (row[j] in registers)
                    // 1/sqrt of the lane's own running diagonal in FP64 (hardware
                    // estimate and one Newton step); at step j lane j's value is
                    // the pivot's, broadcast as the float the rows are scaled by
                    double r = __builtin_amdgcn_rsq(dd);
                    r = r * (1.5 - (0.5 * dd) * (r * r));
                    const float invg = sub_bcast<PPW>(static_cast<float>(r), j, half);
                    if (j < n) {
                        if (li > j && li < n) row[j] *= invg;
                        if (li > j && li < n) lmat[lrow + j] = row[j];
                    }
                    // Branch-free update of the lanes below the pivot (others take a
                    // zero multiplier).  Entries k >= li of row li are its diagonal
                    // (tracked in dd) and upper part, which are never read again, so
                    // they are updated unmasked.  G[k][j] is lane k's row[j]: with
                    // one system per wave it is broadcast straight from that lane
                    // (no LDS round trip per step); two systems per wave read it
                    // back from the factor in LDS.
                    const float nlij = (j < n && li > j && li < n) ? -row[j] : 0.0f;
                    if constexpr (PPW == 1) {
                        float gcol[LMAX];
    #pragma unroll
                        for (int k = j + 1; k < LMAX; k++) gcol[k] = bcast(row[j], k);
    #pragma unroll
                        for (int k = j + 1; k < LMAX; k++) row[k] = fmaf(nlij, gcol[k], row[k]);
                    } else {
                        wave_sync();
    #pragma unroll
                        for (int k = j + 1; k < LMAX; k++) row[k] = fmaf(nlij, lmat[(k * (k - 1)) / 2 + j], row[k]);
                    }
                    dd -= static_cast<double>(nlij) * static_cast<double>(nlij);
                }
            }
            // Row li's diagonal stops changing after step li - 1, so dd is now
            // lane li's pivot: its 1/G[li][li] (the value broadcast at step li)
            // and whether every pivot of the system was positive.
            double rg = __builtin_amdgcn_rsq(dd);
            rg = rg * (1.5 - (0.5 * dd) * (rg * rg));
            const float ginv = static_cast<float>(rg);
            const uint64_t nonpos = __ballot(li < n && !(dd > 0.0));
            const bool ok = (PPW == 1) ? nonpos == 0 : ((nonpos >> (32 * half)) & 0xffffffffull) == 0;
            wave_sync();
            w = chol_solve<LMAX, PPW>(row, ginv, 1.0f, n, li, half, lmat);
            // FP64 residual r = R~ w - 1 and up to two refinement steps, only for
            // systems that may be ill-conditioned.  R~ is a sum of outer products
            // (positive semi-definite), so with the loading delta = alpha trace/n
            // its condition number is at most (trace + delta) / delta = 1 + n/alpha;
            // the ratio of the largest to the smallest squared pivot is a lower
            // bound on it.  Below kKappaSafe (bound) or kPivotRatioSafe (pivots)
            // the FP32 factorisation already solves to about kappa * 6e-8.
            float dmax = (li < n) ? static_cast<float>(dd) : 0.0f;
            float dmin = (li < n) ? static_cast<float>(dd) : FLT_MAX;
            dmax = sub_max<PPW>(dmax);
            dmin = sub_min<PPW>(dmin);
            const bool loaded_ok = a.alpha > 0.0f && 1.0f + static_cast<float>(n) / a.alpha <= kKappaSafe;
            const bool suspect = ok && !loaded_ok && !(dmax <= kPivotRatioSafe * dmin);
            for (int it = 0; it < 3 && __any(suspect); it++) {
                double r = 0.0;
    #pragma unroll
                for (int k = 0; k < LMAX; k++) {
                    const double wk = static_cast<double>(sub_bcast_swz<PPW>(w, k, half));
                    if (k < n) r += static_cast<double>(orig[k]) * wk;
                }
                r = (li < n) ? r - 1.0 : 0.0;
                const double rn = sub_sum<PPW>(r * r);
                // refine only where single precision left a relative residual
                // above 1e-3 (ill-conditioned systems); both halves of a wave
                // take the same number of steps
                const bool more = (rn > 1e-6 * n) && suspect;
                const bool any_more = __any(more);
                if (!any_more || it == 2) break;
                const float dw = chol_solve<LMAX, PPW>(row, ginv, static_cast<float>(r), n, li, half, lmat);
                if (more) w -= dw;
            }
        } else {
            w = pcg_solve<LMAX, PPW>(orig, diag + delta, n, li, half, a.pcg_iters, a.pcg_threshold);
        }
        // weights -> sample (applyWeights, RxBeamformerMV.cu:342-376)
        float xsum = 0.0f;
        if (li < n)
            for (int k = 0; k < m.kact; k++) xsum += xs[ht * rs + m.first + k + li];
        const float wsum = sub_sum<PPW>(li < n ? w : 0.0f);
        const float scaling = 1.0f / (wsum * powf(static_cast<float>(m.kact), a.p));
        float y = sub_sum<PPW>(li < n ? xsum * w * scaling : 0.0f);
        if (fabsf(y) > 1e7f || isnan(y) || bad) y = 0.0f;
        if (li == 0 && valid) out[static_cast<uint64_t>(t) * a.n_rx + s] = clamp_out<OutT>(y * static_cast<float>(a.n_ch));
        wave_sync();
    }
}

// Workgroup b runs on XCD b % 8 (round robin).  Consecutive depth tiles of a
// scanline share their +-T halo of staged samples; renumbered in XCD tiles of
// kMvXcdTile consecutive workgroups (rounds of 8 tiles rotated by one XCD per
// round) they run on one XCD and re-read the halo from its L2.  A bijection
// on the grid, which the launch pads to whole rounds.
constexpr uint32_t kMvXcdTile = 32;   // C3: 4 / 8 / 16 / 32 / 64 / 128 -> 2.045 / 2.04 / 2.035 / 2.025 / 2.02 / 2.06 ms
__device__ inline uint32_t mv_logical_wg() {
    const uint32_t b = blockIdx.x, xcd = b % 8u, local = b / 8u;
    const uint32_t round = local / kMvXcdTile, k = local % kMvXcdTile;
    return (round * 8u + (xcd + round) % 8u) * kMvXcdTile + k;
}

template <typename InT, typename OutT, int LMAX, bool MEANS, bool FUSED>
__global__ __launch_bounds__(kThreads) void mv_kernel(const MvArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t b = mv_logical_wg();
    if (b >= a.n_rx * a.n_tiles) return;   // whole workgroup: the padded grid
    mv_tile<InT, OutT, LMAX, MEANS, FUSED>(a, a.g.lut, b / a.n_tiles, b % a.n_tiles, lds);
}

template <typename InT, typename OutT, int LMAX, bool FUSED>
hipError_t launch_typed(const MvArgs& a, size_t lds, hipStream_t stream) {
    const uint32_t round = 8u * kMvXcdTile;
    dim3 grid((a.n_rx * a.n_tiles + round - 1) / round * round), block(kThreads);
    if (a.means) hipLaunchKernelGGL((mv_kernel<InT, OutT, LMAX, true, FUSED>), grid, block, lds, stream, a);
    else hipLaunchKernelGGL((mv_kernel<InT, OutT, LMAX, false, FUSED>), grid, block, lds, stream, a);
    return hipGetLastError();
}

template <typename InT, typename OutT, bool FUSED>
hipError_t launch_lmax(const MvArgs& a, size_t lds, hipStream_t stream) {
    if (a.L <= 32) return launch_typed<InT, OutT, 32, FUSED>(a, lds, stream);
    return launch_typed<InT, OutT, 64, FUSED>(a, lds, stream);
}

template <bool FUSED>
hipError_t launch_io(const MvArgs& a, int in_dtype, int out_dtype, size_t lds, hipStream_t hs) {
    if (in_dtype == SUPRA_INT16)
        return (out_dtype == SUPRA_FLOAT32) ? launch_lmax<int16_t, float, FUSED>(a, lds, hs)
                                            : launch_lmax<int16_t, int16_t, FUSED>(a, lds, hs);
    return (out_dtype == SUPRA_FLOAT32) ? launch_lmax<float, float, FUSED>(a, lds, hs)
                                        : launch_lmax<float, int16_t, FUSED>(a, lds, hs);
}

// ---- persistent work queue over many frames (fused pre-delay + MV) --------

template <typename InT, int LMAX, bool MEANS>
__global__ __launch_bounds__(kThreads) void mv_batch_kernel(const MvBatchArgs b) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t s_item;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = atomicAdd(b.sync.head, 1u);
        __syncthreads();
        const uint32_t item = s_item;
        if (item >= b.n_items) break;
        const uint32_t code = b.items[item];
        const MvBatchFrame& fr = b.frames[code >> 24];
        const MvArgs a = fr.a;     // scalars of the frame (the LUT is read from fr)
        const uint32_t local = code & 0xffffffu;
        mv_tile<InT, float, LMAX, MEANS, true>(a, fr.a.g.lut, local / a.n_tiles, local % a.n_tiles, lds);
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence();
            if (atomicAdd(&b.sync.done[fr.frame], 1u) + 1u == fr.n_items) b.sync.stamp[fr.frame] = wall_clock64();
        }
    }
}

int max_lds_per_block() {
    static std::mutex mu;
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 65536;
    std::lock_guard<std::mutex> lock(mu);
    if (cached[dev] == 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
            v = 65536;
        cached[dev] = v;
    }
    return cached[dev];
}

// LDS layout for a tile of `tile` output depths; returns total bytes.
size_t plan_lds(MvArgs& a, uint32_t tile) {
    const uint32_t H = tile + 2 * a.T;
    size_t off = static_cast<size_t>(H) * a.rs * 4;
    a.off_meta = static_cast<uint32_t>(off);
    off += static_cast<size_t>(H) * sizeof(SampleMeta);
    a.off_mean = static_cast<uint32_t>(off);
    off += a.means ? static_cast<size_t>(H) * a.L * 4 : 0;
    a.off_r = static_cast<uint32_t>(off);
    off += static_cast<size_t>(H) * a.packed * 4;
    off = (off + 15) & ~static_cast<size_t>(15);
    a.off_wave = static_cast<uint32_t>(off);
    a.wave_bytes = static_cast<uint32_t>(((a.L * (a.L - 1) / 2 * 4) + 15) & ~15u);   // strictly lower factor
    const uint32_t ppw = a.L <= 32 ? 2 : 1;   // pixels per wave in phase 3
    off += static_cast<size_t>(a.waves) * ppw * a.wave_bytes;
    a.off_fused = static_cast<uint32_t>(off);
    if (a.g.rf) off += (kWindowEntries + 4 * static_cast<size_t>(a.n_ch) + 6 * static_cast<size_t>(H)) * 4;
    return off;
}

}  // namespace
}  // namespace supra_amd

using namespace supra_amd;

namespace supra_amd {

// Argument checks and the LDS/tile plan shared by both entry points.
int mv_prepare(MvArgs& a, uint32_t num_scanlines, uint32_t num_channels, uint32_t num_samples,
               uint32_t sub_array_size, uint32_t temporal_smoothing, double sub_array_scaling_power,
               int32_t compute_means, double loading_factor, size_t lds_budget, size_t* lds_out) {
    if (num_scanlines == 0 || num_channels == 0 || num_samples == 0)
        return fail(SUPRA_ERR_INVALID_ARG, "mv: empty input");
    const uint32_t L = sub_array_size == 0 ? num_channels / 2 : sub_array_size;
    if (L == 0 || L > num_channels || L > 64)
        return fail(SUPRA_ERR_UNSUPPORTED, "mv: sub-array size %u must be in [1, min(64, channels)]", L);
    if (num_channels > 256) return fail(SUPRA_ERR_UNSUPPORTED, "mv: at most 256 channels");
    if (temporal_smoothing > 32) return fail(SUPRA_ERR_UNSUPPORTED, "mv: temporal smoothing above 32");
    if (!(loading_factor >= 0) || !std::isfinite(sub_array_scaling_power))
        return fail(SUPRA_ERR_INVALID_ARG, "mv: loading factor must be >= 0");
    a.n_rx = num_scanlines;
    a.n_ch = num_channels;
    a.n_d = num_samples;
    a.L = L;
    a.T = temporal_smoothing;
    a.K = num_channels - L + 1;
    a.rs = num_channels + L + 8;   // zero padding for the sliding windows of the covariance phase
    a.packed = L * (L + 1) / 2;
    a.p = static_cast<float>(sub_array_scaling_power);
    a.alpha = static_cast<float>(loading_factor);
    a.means = compute_means ? 1 : 0;
    const size_t budget = lds_budget ? std::min(lds_budget, static_cast<size_t>(max_lds_per_block()))
                                     : static_cast<size_t>(max_lds_per_block());
    // Without an explicit budget, prefer two workgroups per CU (the solve
    // phase is latency-bound; a second resident workgroup hides it) as long as
    // the tile keeps the temporal halo overhead moderate; else one workgroup
    // with the largest tile.
    uint32_t tile = 0;
    const uint32_t min_tile_half = lds_budget ? kMaxTile + 1 : std::max<uint32_t>(8, 2 * a.T + 2);
    a.waves = kWaves;
    for (uint32_t dt = kMaxTile; dt >= min_tile_half; dt--) {
        a.tile = dt;
        if (plan_lds(a, dt) <= budget / 2) {
            tile = dt;
            break;
        }
    }
    for (uint32_t waves = kWaves; waves >= 1 && tile == 0; waves /= 2) {
        a.waves = waves;
        for (uint32_t dt = kMaxTile; dt >= 1; dt--) {
            a.tile = dt;
            if (plan_lds(a, dt) <= budget) {
                tile = dt;
                break;
            }
        }
    }
    if (tile == 0) return fail(SUPRA_ERR_TOO_LARGE, "mv: configuration does not fit the %zu-byte LDS", budget);
    a.tile = std::min<uint32_t>(tile, num_samples);
    *lds_out = plan_lds(a, a.tile);
    a.n_tiles = (num_samples + a.tile - 1) / a.tile;
    if (static_cast<uint64_t>(a.n_tiles) * num_scanlines >= (1ull << 31))
        return fail(SUPRA_ERR_TOO_LARGE, "mv: too many workgroups");
    return SUPRA_OK;
}

hipError_t launch_mv_batch(const MvBatchArgs& b, uint32_t lmax, int32_t rf_dtype, int32_t means, uint32_t workgroups,
                           size_t lds, hipStream_t stream) {
    const dim3 grid(workgroups), block(kThreads);
#define SUPRA_MV_BATCH(T, LM, M) hipLaunchKernelGGL((mv_batch_kernel<T, LM, M>), grid, block, lds, stream, b)
    if (rf_dtype == SUPRA_INT16) {
        if (lmax <= 32) { if (means) SUPRA_MV_BATCH(int16_t, 32, true); else SUPRA_MV_BATCH(int16_t, 32, false); }
        else { if (means) SUPRA_MV_BATCH(int16_t, 64, true); else SUPRA_MV_BATCH(int16_t, 64, false); }
    } else {
        if (lmax <= 32) { if (means) SUPRA_MV_BATCH(float, 32, true); else SUPRA_MV_BATCH(float, 32, false); }
        else { if (means) SUPRA_MV_BATCH(float, 64, true); else SUPRA_MV_BATCH(float, 64, false); }
    }
#undef SUPRA_MV_BATCH
    return hipGetLastError();
}

size_t mv_batch_frame_bytes() { return sizeof(MvBatchFrame); }

int launch_mv_fused(const DasArgs& g, int32_t rf_dtype, uint32_t sub_array_size, uint32_t temporal_smoothing,
                    double sub_array_scaling_power, int32_t compute_means, double loading_factor, void* out,
                    int32_t out_dtype, hipStream_t stream) {
    MvArgs a;
    std::memset(&a, 0, sizeof(a));
    a.g = g;
    a.out = out;
    size_t lds = 0;
    int st = mv_prepare(a, g.n_rx, g.n_ch, g.n_depths, sub_array_size, temporal_smoothing, sub_array_scaling_power,
                        compute_means, loading_factor, 0, &lds);
    if (st != SUPRA_OK) return st;
    hipError_t e = launch_io<true>(a, rf_dtype, out_dtype, lds, stream);
    if (e != hipSuccess) return fail(SUPRA_ERR_HIP, "mv: launch: %s", hipGetErrorString(e));
    return SUPRA_OK;
}

}  // namespace supra_amd

extern "C" int supra_mv_run(const void* raw_delayed, int32_t in_dtype, uint32_t num_scanlines, uint32_t num_channels,
                            uint32_t num_samples, uint32_t sub_array_size, uint32_t temporal_smoothing,
                            double sub_array_scaling_power, int32_t compute_means, double loading_factor, void* out,
                            int32_t out_dtype, void* stream) {
    clear_error();
    if (!raw_delayed || !out) return fail(SUPRA_ERR_INVALID_ARG, "mv: null argument");
    if ((in_dtype != SUPRA_INT16 && in_dtype != SUPRA_FLOAT32) ||
        (out_dtype != SUPRA_INT16 && out_dtype != SUPRA_FLOAT32))
        return fail(SUPRA_ERR_UNSUPPORTED, "mv: element types must be SUPRA_INT16 or SUPRA_FLOAT32");
    MvArgs a;
    std::memset(&a, 0, sizeof(a));
    a.x = raw_delayed;
    a.out = out;
    size_t lds = 0;
    int st = mv_prepare(a, num_scanlines, num_channels, num_samples, sub_array_size, temporal_smoothing,
                        sub_array_scaling_power, compute_means, loading_factor, 0, &lds);
    if (st != SUPRA_OK) return st;
    hipError_t e = launch_io<false>(a, in_dtype, out_dtype, lds, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail(SUPRA_ERR_HIP, "mv: launch: %s", hipGetErrorString(e));
    return SUPRA_OK;
}

extern "C" int supra_mvpcg_run(const void* raw_delayed, int32_t in_dtype, uint32_t num_scanlines,
                               uint32_t num_channels, uint32_t num_samples, uint32_t sub_array_size,
                               uint32_t temporal_smoothing, uint32_t max_iterations_override,
                               double convergence_threshold, double sub_array_scaling_power, int32_t compute_means,
                               void* out, int32_t out_dtype, void* stream) {
    clear_error();
    if (!raw_delayed || !out) return fail(SUPRA_ERR_INVALID_ARG, "mvpcg: null argument");
    if ((in_dtype != SUPRA_INT16 && in_dtype != SUPRA_FLOAT32) ||
        (out_dtype != SUPRA_INT16 && out_dtype != SUPRA_FLOAT32))
        return fail(SUPRA_ERR_UNSUPPORTED, "mvpcg: element types must be SUPRA_INT16 or SUPRA_FLOAT32");
    if (!(convergence_threshold >= 0) || !std::isfinite(convergence_threshold))
        return fail(SUPRA_ERR_INVALID_ARG, "mvpcg: convergence threshold %g", convergence_threshold);
    if (max_iterations_override > 10000)
        return fail(SUPRA_ERR_INVALID_ARG, "mvpcg: at most 10000 iterations (%u)", max_iterations_override);
    MvArgs a;
    std::memset(&a, 0, sizeof(a));
    a.x = raw_delayed;
    a.out = out;
    size_t lds = 0;
    // loading trace / L_t as addDiagonalLoading (RxBeamformerMVpcg.cu:295-330)
    int st = mv_prepare(a, num_scanlines, num_channels, num_samples, sub_array_size, temporal_smoothing,
                        sub_array_scaling_power, compute_means, 1.0, 0, &lds);
    if (st != SUPRA_OK) return st;
    // maxIterations = the sub-array size unless overridden (:649)
    a.pcg_iters = max_iterations_override ? max_iterations_override : a.L;
    a.pcg_threshold = static_cast<float>(convergence_threshold);
    hipError_t e = launch_io<false>(a, in_dtype, out_dtype, lds, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail(SUPRA_ERR_HIP, "mvpcg: launch: %s", hipGetErrorString(e));
    return SUPRA_OK;
}
