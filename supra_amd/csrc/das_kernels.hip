// Receive delay-and-sum kernels for gfx950 (2D linear/phased scans and 3D
// matrix-probe volumes).
//
// Math follows the reference's per-sample beamformer
// (src/SupraLib/Beamformer/RxSampleBeamformerDelayAndSum.h:25-208), the
// delay/aperture helpers (RxBeamformerCommon.h:21-61), the window lookup
// (WindowFunction.h:58-74) and the per-pixel driver kernels
// (RxBeamformerCuda.cu:109-323).  Every single-precision operation that feeds
// a discontinuity (aperture membership, floor of the delay, window index) is
// issued in the reference's operand order with IEEE rounding: this file is
// compiled with -ffp-contract=off and without fast-math, and sqrt/division
// are correctly rounded, so the kernels reproduce the host restatement bit
// for bit, including the summation order over the aperture.
//
// Launch layout (MI355X): one workgroup of 256 lanes (4 wave64) per
// (frame, rx scanline, tile of 256 depths).  Lanes map to consecutive
// depths, so the two RF gathers of one element hit a contiguous window of
// ~2x256 samples of one channel (coalesced into a few cache lines).  The
// per-element quantities that do not depend on depth (element position
// relative to the scanline, channel offset, their squares) are computed once
// per workgroup into LDS and broadcast to all lanes.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "bf_internal.hpp"
#include "das_launch.hpp"

namespace supra_amd {

namespace {

constexpr int kThreads = 256;
constexpr int kChunk2D = 512;     // aperture elements staged per LDS pass (2D)
constexpr int kChunk3D = 1024;    // aperture elements staged per LDS pass (3D)
constexpr uint32_t kNotConnected = 0xffffffffu;

// clampCast (src/SupraLib/utilities/cudaUtility.h:111-147): fmaxf/fminf keep
// the non-NaN operand, so a NaN input lands on the lower limit.
template <typename T> __device__ inline T clamp_cast(float x);
template <> __device__ inline float clamp_cast<float>(float x) {
    return fminf(fmaxf(x, -FLT_MAX), FLT_MAX);
}
template <> __device__ inline int16_t clamp_cast<int16_t>(float x) {
    return static_cast<int16_t>(fminf(fmaxf(x, -32767.0f), 32767.0f));
}

template <typename T> __device__ inline float to_f(T v) { return static_cast<float>(v); }

// WindowFunctionGpu::get (WindowFunction.h:58-64) on an LDS copy of the LUT.
__device__ inline int window_index(float rel) {
    float cl = fminf(fmaxf(rel, -1.0f), 1.0f);
    return static_cast<int>(static_cast<uint32_t>(roundf(31.5f * (cl + 1.0f))));
}

// Correctly rounded single-precision square root for x == 0 or normal x
// (every argument here is a sum of squares of dt-space lengths, so it is
// either exactly zero or far above 2^-96).  v_sqrt_f32 is within one ulp;
// one residual test on each neighbour picks the correctly rounded value
// (the same correction LLVM emits for sqrtf, without the denormal rescale
// and the class test that this domain does not need).
__device__ inline float sqrt_rn(float x) {
    float s = __builtin_amdgcn_sqrtf(x);
    uint32_t si = __float_as_uint(s);
    float down = __uint_as_float(si - 1u);
    float up = __uint_as_float(si + 1u);
    float r_down = fmaf(-down, s, x);
    float r_up = fmaf(-up, s, x);
    s = (r_down <= 0.0f) ? down : s;
    s = (r_up > 0.0f) ? up : s;
    return s;
}

struct Pixel2D {
    float d, aDT, inv, px, py2;
};

// One transmit's contribution for a 2D pixel, elements [e0, e1) of the
// current LDS chunk (RxSampleBeamformerDelayAndSum.h:122-208).
template <typename RFT>
__device__ inline void accumulate2d(const RFT* __restrict__ rf, uint32_t n_s, float t0, float offset,
                                    const Pixel2D& p, bool vertical, const float* s_xe, const float* s_dx,
                                    const float* s_dx2, const uint32_t* s_off, const float* s_lut, int cnt,
                                    float& sample, float& wsum, int& nadd) {
    for (int j = 0; j < cnt; j++) {
        uint32_t off = s_off[j];
        if (off == kNotConnected) continue;
        float dx = s_dx[j];
        if (fabsf(dx) <= p.aDT) {
            float w = s_lut[window_index(dx * p.inv)];
            wsum += w;
            nadd++;
            float a2;
            if (vertical) {
                a2 = s_dx2[j];
            } else {
                float a = s_xe[j] - p.px;
                a2 = a * a;
            }
            float tau = t0 + (sqrt_rn(a2 + p.py2) + p.d);
            tau = tau + offset;
            float fl = floorf(tau);
            int32_t i = static_cast<int32_t>(fl);
            float f = tau - static_cast<float>(i);
            if (static_cast<uint32_t>(i) < n_s - 1) {
                float v0 = to_f(rf[off + static_cast<uint32_t>(i)]);
                float v1 = to_f(rf[off + static_cast<uint32_t>(i) + 1]);
                sample += w * ((1.0f - f) * v0 + f * v1);
            } else if (static_cast<uint32_t>(i) < n_s && f == 0.0f) {
                sample += w * to_f(rf[off + static_cast<uint32_t>(i)]);
            }
        }
    }
}

// RxSampleBeamformerTestSignal::sampleBeamform2D (RxSampleBeamformerTestSignal.h:87-133):
// 1000 inside 2 mm cylinders every 6 mm at 30 mm depth, 0 elsewhere.
__device__ inline float test_signal_2d(float x, float dir_x, float dir_y, float dir_z, float d, float dt, float c) {
    float px = ((x + dir_x * d) * dt) * c;
    float py = ((dir_y * d) * dt) * c;
    float sample = 0.0f;
    for (int k = -3; k <= 3; k++) {
        float ex = px - static_cast<float>(k) * 6.0f;
        float ey = py - 30.0f;
        float ez = 0.0f;
        float dist = sqrtf(ex * ex + ey * ey + ez * ez);
        if (dist <= 2.0f) sample = 1000.0f;
    }
    return sample;
}

template <typename RFT, typename OutT, bool TESTSIG, bool INTERP_TX, bool MAP>
__global__ __launch_bounds__(kThreads) void das2d_kernel(const DasArgs a) {
    __shared__ float s_lut[kWindowEntries];
    __shared__ float s_xe[kChunk2D];
    __shared__ float s_dx[kChunk2D];
    __shared__ float s_dx2[kChunk2D];
    __shared__ uint32_t s_off[kChunk2D];

    const uint32_t bid = blockIdx.x;
    const uint32_t tile = bid % a.depth_tiles;
    const uint32_t rest = bid / a.depth_tiles;
    const uint32_t s = rest % a.n_rx;
    const uint32_t frame = rest / a.n_rx;
    const uint32_t r = tile * kThreads + threadIdx.x;
    const bool valid = r < a.n_depths;

    for (int i = threadIdx.x; i < kWindowEntries; i += kThreads) s_lut[i] = a.lut[i];

    const ScanlineDT sc = a.scans[s];
    Pixel2D p;
    p.d = valid ? a.depths[r] : 0.0f;
    p.aDT = (((p.d * a.dt) * a.c) / (2.0f * a.F)) / a.c / a.dt;
    p.inv = 1.0f / fminf(p.aDT, sc.max_ed_x);
    p.px = sc.x + sc.dir_x * p.d;
    float py = sc.dir_y * p.d;
    p.py2 = py * py;
    const bool vertical = (sc.dir_x == 0.0f);

    const RFT* rf_frame = static_cast<const RFT*>(a.rf) + static_cast<uint64_t>(frame) * a.rf_frame_stride;
    float s_interp = 0.0f;

    for (int k = 0; k < 4; k++) {
        const bool use = INTERP_TX ? ((sc.active_mask >> k) & 1) != 0 : (k == sc.sel);
        if (!use) continue;   // uniform over the workgroup
        const TxDT tx = sc.tx[k];
        float s_local;
        if (TESTSIG) {
            s_local = test_signal_2d(sc.x, sc.dir_x, sc.dir_y, sc.dir_z, p.d, a.dt, a.c);
        } else {
            const RFT* rf_tx = rf_frame + static_cast<uint64_t>(tx.tx_idx) * a.n_ch * a.n_s;
            float sample = 0.0f, wsum = 0.0f;
            int nadd = 0;
            const int first = tx.first_x, last = tx.last_x;
            for (int base = first; base < last; base += kChunk2D) {
                const int cnt = min(kChunk2D, last - base);
                __syncthreads();
                for (int j = threadIdx.x; j < cnt; j += kThreads) {
                    const int e = base + j;
                    uint32_t ch;
                    if (MAP) {
                        int32_t m = a.e2c[e];
                        ch = (m < 0) ? kNotConnected : static_cast<uint32_t>(m);
                    } else {
                        ch = static_cast<uint32_t>(e) % a.n_ch;
                    }
                    const float xe = a.ex[e];
                    const float dx = xe - sc.x;
                    s_xe[j] = xe;
                    s_dx[j] = dx;
                    s_dx2[j] = dx * dx;
                    s_off[j] = (ch == kNotConnected) ? kNotConnected : ch * a.n_s;
                }
                __syncthreads();
                accumulate2d<RFT>(rf_tx, a.n_s, tx.t0, a.offset, p, vertical, s_xe, s_dx, s_dx2, s_off, s_lut,
                                  cnt, sample, wsum, nadd);
            }
            s_local = (nadd > 0) ? (sample / wsum) * static_cast<float>(nadd) : 0.0f;
        }
        if (INTERP_TX) s_interp += sc.txw[k] * s_local;
        else s_interp += s_local;
    }
    if (valid) {
        OutT* out = static_cast<OutT*>(a.out) + static_cast<uint64_t>(frame) * a.out_frame_stride;
        out[static_cast<uint64_t>(r) * a.n_rx + s] = clamp_cast<OutT>(s_interp);
    }
}

// ---- 3D ---------------------------------------------------------------------

struct Pixel3D {
    float d, aDT, invx, invz, px, pz, py2;
};

// RxSampleBeamformerTestSignal::sampleBeamform3D (RxSampleBeamformerTestSignal.h:25-85).
__device__ inline float test_signal_3d(const ScanlineDT& sc, float d, float dt, float c) {
    float px = ((sc.x + sc.dir_x * d) * dt) * c;
    float py = ((sc.dir_y * d) * dt) * c;
    float pz = ((sc.z + sc.dir_z * d) * dt) * c;
    float sample = 0.0f;
    for (int k = -3; k <= 3; k++) {
        float ex = px - static_cast<float>(k) * 6.0f, ey = py - 30.0f, ez = 0.0f;
        if (sqrtf(ex * ex + ey * ey + ez * ez) <= 2.0f) sample = 1000.0f;
    }
    for (int k = -3; k <= 3; k++) {
        float ex = 0.0f, ey = py - 30.0f, ez = pz - static_cast<float>(k) * 6.0f;
        if (sqrtf(ex * ex + ey * ey + ez * ez) <= 2.0f) sample = 1000.0f;
    }
    return sample;
}

template <typename RFT, typename OutT, bool TESTSIG, bool INTERP_TX, bool MAP>
__global__ __launch_bounds__(kThreads) void das3d_kernel(const DasArgs a) {
    __shared__ float s_w2[kWindowEntries * kWindowEntries];   // sqrt(LUT[i] * LUT[j])
    __shared__ float s_xe[kChunk3D];
    __shared__ float s_ze[kChunk3D];
    __shared__ float s_r2[kChunk3D];
    __shared__ uint32_t s_off[kChunk3D];

    const uint32_t bid = blockIdx.x;
    const uint32_t tile = bid % a.depth_tiles;
    const uint32_t rest = bid / a.depth_tiles;
    const uint32_t s = rest % a.n_rx;
    const uint32_t frame = rest / a.n_rx;
    const uint32_t r = tile * kThreads + threadIdx.x;
    const bool valid = r < a.n_depths;

    // computeWindow3DShared (RxBeamformerCommon.h:54-61) as a 64x64 table:
    // the product and the square root are the same IEEE operations.
    for (int i = threadIdx.x; i < kWindowEntries * kWindowEntries; i += kThreads)
        s_w2[i] = sqrtf(a.lut[i / kWindowEntries] * a.lut[i % kWindowEntries]);

    const ScanlineDT sc = a.scans[s];
    Pixel3D p;
    p.d = valid ? a.depths[r] : 0.0f;
    {
        float t = (((p.d * a.dt) * a.c) / (2.0f * a.F)) / a.c / a.dt;
        p.aDT = t * t;
        float sq = sqrtf(p.aDT);
        p.invx = 1.0f / fminf(sq, sc.max_ed_x);
        p.invz = 1.0f / fminf(sq, sc.max_ed_y);
    }
    p.px = sc.x + sc.dir_x * p.d;
    p.pz = sc.z + sc.dir_z * p.d;
    {
        float py = sc.dir_y * p.d;
        p.py2 = py * py;
    }

    const RFT* rf_frame = static_cast<const RFT*>(a.rf) + static_cast<uint64_t>(frame) * a.rf_frame_stride;
    float s_interp = 0.0f;

    for (int k = 0; k < 4; k++) {
        const bool use = INTERP_TX ? ((sc.active_mask >> k) & 1) != 0 : (k == sc.sel);
        if (!use) continue;
        const TxDT tx = sc.tx[k];
        float s_local;
        if (TESTSIG) {
            s_local = test_signal_3d(sc, p.d, a.dt, a.c);
        } else {
            const RFT* rf_tx = rf_frame + static_cast<uint64_t>(tx.tx_idx) * a.n_ch * a.n_s;
            float sample = 0.0f, wsum = 0.0f;
            int nadd = 0;
            const int ny = (tx.last_y > tx.first_y) ? tx.last_y - tx.first_y : 0;
            const int nx = (tx.last_x > tx.first_x) ? tx.last_x - tx.first_x : 0;
            const int total = nx * ny;
            // element j of the aperture: ex = first_x + j / ny, ey = first_y + j % ny
            // (x outer, y inner: RxSampleBeamformerDelayAndSum.h:55-59)
            for (int base = 0; base < total; base += kChunk3D) {
                const int cnt = min(kChunk3D, total - base);
                __syncthreads();
                for (int j = threadIdx.x; j < cnt; j += kThreads) {
                    const int jj = base + j;
                    const uint32_t exi = tx.first_x + jj / ny;
                    const uint32_t eyi = tx.first_y + jj % ny;
                    const uint32_t e = exi + eyi * a.layout_x;
                    uint32_t ch;
                    if (MAP) {
                        int32_t m = a.e2c[e];
                        ch = (m < 0) ? kNotConnected : static_cast<uint32_t>(m);
                    } else {
                        ch = e % a.n_ch;
                    }
                    const float xe = a.ex[e], ze = a.ey[e];
                    const float dx = xe - sc.x, dz = ze - sc.z;
                    s_xe[j] = xe;
                    s_ze[j] = ze;
                    s_r2[j] = dx * dx + dz * dz;
                    s_off[j] = (ch == kNotConnected) ? kNotConnected : ch * a.n_s;
                }
                __syncthreads();
                for (int j = 0; j < cnt; j++) {
                    const uint32_t off = s_off[j];
                    if (off == kNotConnected) continue;
                    if (s_r2[j] <= p.aDT) {
                        const float xe = s_xe[j], ze = s_ze[j];
                        const int ix = window_index((xe - sc.x) * p.invx);
                        const int iz = window_index((ze - sc.z) * p.invz);
                        const float w = s_w2[ix * kWindowEntries + iz];
                        wsum += w;
                        nadd++;
                        const float ax = xe - p.px, az = ze - p.pz;
                        float tau = tx.t0 + (sqrt_rn((ax * ax + az * az) + p.py2) + p.d);
                        tau = tau + a.offset;
                        // floor then a saturating float->uint32 conversion, as the
                        // reference's static_cast<uint32_t>(floor(.)) on the GPU
                        // (RxSampleBeamformerDelayAndSum.h:87-88): a negative delay
                        // becomes sample 0 with a negative fraction.
                        const float fl = floorf(tau);
                        const uint32_t i = (fl > 0.0f) ? static_cast<uint32_t>(fl) : 0u;
                        const float f = tau - static_cast<float>(i);
                        if (i < a.n_s - 1) {
                            const float v0 = to_f(rf_tx[off + i]);
                            const float v1 = to_f(rf_tx[off + i + 1]);
                            sample += w * ((1.0f - f) * v0 + f * v1);
                        } else if (i < a.n_s && f == 0.0f) {
                            sample += w * to_f(rf_tx[off + i]);
                        }
                    }
                }
            }
            s_local = (nadd > 0) ? (sample / wsum) * static_cast<float>(nadd) : 0.0f;
        }
        if (INTERP_TX) s_interp += sc.txw[k] * s_local;
        else s_interp += s_local;
    }
    if (valid) {
        OutT* out = static_cast<OutT*>(a.out) + static_cast<uint64_t>(frame) * a.out_frame_stride;
        out[static_cast<uint64_t>(r) * a.n_rx + s] = clamp_cast<OutT>(s_interp);
    }
}

}  // namespace

template <typename RFT, typename OutT>
static hipError_t launch_typed(const DasArgs& a, bool is3d, bool testsig, bool interp_tx, bool map,
                               hipStream_t stream) {
    const uint64_t blocks = static_cast<uint64_t>(a.depth_tiles) * a.n_rx * a.n_frames;
    dim3 grid(static_cast<uint32_t>(blocks)), block(kThreads);
#define SUPRA_DAS_DISPATCH(KERNEL)                                                              \
    do {                                                                                         \
        if (testsig) {                                                                           \
            if (interp_tx) hipLaunchKernelGGL((KERNEL<RFT, OutT, true, true, false>), grid, block, 0, stream, a);   \
            else hipLaunchKernelGGL((KERNEL<RFT, OutT, true, false, false>), grid, block, 0, stream, a);            \
        } else if (map) {                                                                        \
            if (interp_tx) hipLaunchKernelGGL((KERNEL<RFT, OutT, false, true, true>), grid, block, 0, stream, a);   \
            else hipLaunchKernelGGL((KERNEL<RFT, OutT, false, false, true>), grid, block, 0, stream, a);            \
        } else {                                                                                 \
            if (interp_tx) hipLaunchKernelGGL((KERNEL<RFT, OutT, false, true, false>), grid, block, 0, stream, a);  \
            else hipLaunchKernelGGL((KERNEL<RFT, OutT, false, false, false>), grid, block, 0, stream, a);           \
        }                                                                                        \
    } while (0)
    if (is3d) SUPRA_DAS_DISPATCH(das3d_kernel);
    else SUPRA_DAS_DISPATCH(das2d_kernel);
#undef SUPRA_DAS_DISPATCH
    return hipGetLastError();
}

uint32_t das_depth_tile() { return kThreads; }

hipError_t launch_das(const DasArgs& a, int rf_dtype, int out_dtype, bool is3d, bool testsig, bool interp_tx,
                      bool map, hipStream_t stream) {
    if (rf_dtype == SUPRA_INT16) {
        if (out_dtype == SUPRA_FLOAT32) return launch_typed<int16_t, float>(a, is3d, testsig, interp_tx, map, stream);
        return launch_typed<int16_t, int16_t>(a, is3d, testsig, interp_tx, map, stream);
    }
    if (out_dtype == SUPRA_FLOAT32) return launch_typed<float, float>(a, is3d, testsig, interp_tx, map, stream);
    return launch_typed<float, int16_t>(a, is3d, testsig, interp_tx, map, stream);
}

}  // namespace supra_amd
