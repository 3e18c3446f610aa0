// The two other receive beamformers of SUPRA's BeamformingNode
// (beamformerType DelayAndStdDev and CoherenceFactorDelayAndSum), 2D and 3D:
//   RxSampleBeamformerDelayAndStdDev.h:36-158 (3D), :161-270 (2D)
//   RxSampleBeamformerCoherenceFactorDelayAndSum.h:33-150 (3D), :152-258 (2D)
// driven per pixel like the delay-and-sum kernels (RxBeamformerCuda.cu:109-323).
//
// Both make two passes over the aperture: the first is the delay-and-sum of
// the pixel (the same single-precision operations, in the same order, as
// das_kernels.hip, so its value is bit for bit the DelayAndSum image), the
// second accumulates
//   DelayAndStdDev:  sd += (w v - das)^2,   result sqrt(sd) / W * n
//   CoherenceFactor: cs += v, te += v^2,    result cs^2 / te / n * das
// over the aperture members, v the interpolated sample.  A member whose
// delay is out of range has v = 0 (the reference keeps the previous
// element's value in that case).  Lanes are consecutive depths of one
// scanline; element positions come through scalar loads.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "bf_internal.hpp"
#include "das_device.hpp"
#include "das_launch.hpp"

namespace supra_amd {
namespace {

constexpr int kThreads = 256;
constexpr uint32_t kNoChannel = 0xffffffffu;

// The interpolated sample at delay tau with the reference's range rules (see
// gather() in das_device.hpp), or 0 outside them.
template <typename RFT, bool SAT_UINT>
__device__ inline float interp_at(__amdgpu_buffer_rsrc_t rf, uint32_t off, uint32_t n_s, float tau) {
    uint32_t ui;
    float f;
    if constexpr (SAT_UINT) {
        const float fl = fmaxf(floorf(tau), 0.0f);
        ui = static_cast<uint32_t>(fl);
        f = tau - fl;
    } else {
        ui = static_cast<uint32_t>(static_cast<int32_t>(floorf(tau)));
        f = __builtin_amdgcn_fractf(tau);
    }
    const float v0 = load_rf<RFT>(rf, off + ui);
    const float v1 = load_rf<RFT>(rf, off + ui + 1u);
    if constexpr (std::is_same_v<RFT, int16_t>) {
        const float last = static_cast<float>(n_s - 1u);
        bool in;
        if constexpr (SAT_UINT) in = !(tau > last) & !(tau < (n_s > 1u ? -INFINITY : 0.0f));
        else in = (tau >= 0.0f) & (tau <= last);
        return in ? ((1.0f - f) * v0 + f * v1) : 0.0f;
    } else {
        const bool inr = ui < n_s - 1u;
        const bool edge = (ui == n_s - 1u) & (f == 0.0f);
        return inr ? ((1.0f - f) * v0 + f * v1) : (edge ? v0 : 0.0f);
    }
}

__device__ inline uint32_t channel_offset(const DasArgs& a, uint32_t e) {
    if (a.e2c) {
        const int32_t m = a.e2c[e];
        return m < 0 ? kNoChannel : static_cast<uint32_t>(m) * a.n_s;
    }
    return (e % a.n_ch) * a.n_s;
}

// Combines the two passes into the beamformer's per-transmit value.
template <int KIND>
struct Second {
    float sd = 0.0f, cs = 0.0f, te = 0.0f;
    __device__ void add(float w, float v, float das) {
        if (KIND == SUPRA_BF_DELAY_AND_STD_DEV) {
            const float dv = w * v - das;
            sd += dv * dv;
        } else {
            cs += v;
            te += v * v;
        }
    }
    __device__ float result(float das, float wsum, int nadd) const {
        if (nadd <= 0) return 0.0f;
        if (KIND == SUPRA_BF_DELAY_AND_STD_DEV) return sqrtf(sd) / wsum * static_cast<float>(nadd);
        return cs * cs / te / static_cast<float>(nadd) * das;
    }
};

template <typename RFT, typename OutT, int KIND, bool INTERP_TX>
__global__ __launch_bounds__(kThreads) void das2d_variant_kernel(const DasArgs a) {
    __shared__ float s_lut[kWindowEntries];
    const uint32_t bid = blockIdx.x;
    const uint32_t s = (bid / a.depth_tiles) % a.n_rx;
    const uint32_t frame = (bid / a.depth_tiles) / a.n_rx;
    const uint32_t r = (bid % a.depth_tiles) * kThreads + threadIdx.x;
    const bool valid = r < a.n_depths;
    for (int i = threadIdx.x; i < kWindowEntries; i += kThreads) s_lut[i] = a.lut[i];
    __syncthreads();
    const ScanView sc = scan_view(a.scans, s);
    const float d = valid ? a.depths[r] : 0.0f;
    const float aDT = (((d * a.dt) * a.c) / (2.0f * a.F)) / a.c / a.dt;
    const float inv = 1.0f / fminf(aDT, sc.max_ed_x());
    const float x = sc.x();
    const float px = x + sc.dir_x() * d;
    const float py = sc.dir_y() * d;
    const float py2 = py * py;
    const RFT* rf_frame = static_cast<const RFT*>(a.rf) + static_cast<uint64_t>(frame) * a.rf_frame_stride;
    float s_interp = 0.0f;
    for (int k = 0; k < 4; k++) {
        const bool use = INTERP_TX ? ((sc.active_mask() >> k) & 1) != 0 : (k == sc.sel());
        if (!use) continue;
        const TxDT tx = sc.tx(k);
        const RFT* rf_tx = rf_frame + static_cast<uint64_t>(tx.tx_idx) * a.n_ch * a.n_s;
        const __amdgpu_buffer_rsrc_t rsrc = rf_resource(rf_tx, a.n_ch * a.n_s * sizeof(RFT));
        float sample = 0.0f, wsum = 0.0f;
        int nadd = 0;
        for (int pass = 0; pass < 2; pass++) {
            Second<KIND> acc;
            const float das = (nadd > 0) ? (sample / wsum) * static_cast<float>(nadd) : 0.0f;
            for (uint32_t e = tx.first_x; e < tx.last_x; e++) {
                const uint32_t off = channel_offset(a, e);
                if (off == kNoChannel) continue;
                const float xe = load_uniform_float(a.ex, e);
                const float dx = xe - x;
                if (!(fabsf(dx) <= aDT)) continue;
                const float w = s_lut[window_index(dx * inv)];
                const float ex = xe - px;
                float tau = tx.t0 + (sqrt_rn(ex * ex + py2) + d);
                tau = tau + a.offset;
                const float v = interp_at<RFT, false>(rsrc, off, a.n_s, tau);
                if (pass == 0) {
                    wsum += w;
                    nadd++;
                    sample += w * v;
                } else {
                    acc.add(w, v, das);
                }
            }
            if (pass == 1) {
                const float s_local = acc.result(das, wsum, nadd);
                if (INTERP_TX) s_interp += sc.txw(k) * s_local;
                else s_interp += s_local;
            }
        }
    }
    if (valid) {
        OutT* out = static_cast<OutT*>(a.out) + static_cast<uint64_t>(frame) * a.out_frame_stride;
        out[static_cast<uint64_t>(r) * a.n_rx + s] = clamp_cast<OutT>(s_interp);
    }
}

template <typename RFT, typename OutT, int KIND, bool INTERP_TX>
__global__ __launch_bounds__(kThreads) void das3d_variant_kernel(const DasArgs a) {
    __shared__ float s_lut[kWindowEntries];
    const uint32_t bid = blockIdx.x;
    const uint32_t s = (bid / a.depth_tiles) % a.n_rx;
    const uint32_t frame = (bid / a.depth_tiles) / a.n_rx;
    const uint32_t r = (bid % a.depth_tiles) * kThreads + threadIdx.x;
    const bool valid = r < a.n_depths;
    for (int i = threadIdx.x; i < kWindowEntries; i += kThreads) s_lut[i] = a.lut[i];
    __syncthreads();
    const ScanView sc = scan_view(a.scans, s);
    const float d = valid ? a.depths[r] : 0.0f;
    const float ap = (((d * a.dt) * a.c) / (2.0f * a.F)) / a.c / a.dt;
    const float aDT = ap * ap;
    const float sq = sqrtf(aDT);
    const float invx = 1.0f / fminf(sq, sc.max_ed_x());
    const float invz = 1.0f / fminf(sq, sc.max_ed_y());
    const float x = sc.x(), z = sc.z();
    const float px = x + sc.dir_x() * d;
    const float pz = z + sc.dir_z() * d;
    const float py = sc.dir_y() * d;
    const float py2 = py * py;
    const RFT* rf_frame = static_cast<const RFT*>(a.rf) + static_cast<uint64_t>(frame) * a.rf_frame_stride;
    float s_interp = 0.0f;
    for (int k = 0; k < 4; k++) {
        const bool use = INTERP_TX ? ((sc.active_mask() >> k) & 1) != 0 : (k == sc.sel());
        if (!use) continue;
        const TxDT tx = sc.tx(k);
        const RFT* rf_tx = rf_frame + static_cast<uint64_t>(tx.tx_idx) * a.n_ch * a.n_s;
        const __amdgpu_buffer_rsrc_t rsrc = rf_resource(rf_tx, a.n_ch * a.n_s * sizeof(RFT));
        float sample = 0.0f, wsum = 0.0f;
        int nadd = 0;
        for (int pass = 0; pass < 2; pass++) {
            Second<KIND> acc;
            const float das = (nadd > 0) ? (sample / wsum) * static_cast<float>(nadd) : 0.0f;
            for (uint32_t exi = tx.first_x; exi < tx.last_x; exi++) {
                for (uint32_t eyi = tx.first_y; eyi < tx.last_y; eyi++) {
                    const uint32_t e = exi + eyi * a.layout_x;
                    const uint32_t off = channel_offset(a, e);
                    if (off == kNoChannel) continue;
                    const float xe = load_uniform_float(a.ex, e), ze = load_uniform_float(a.ey, e);
                    const float dx = xe - x, dz = ze - z;
                    if (!((dx * dx + dz * dz) <= aDT)) continue;
                    const float w = sqrtf(s_lut[window_index(dx * invx)] * s_lut[window_index(dz * invz)]);
                    const float ax = xe - px, az = ze - pz;
                    float tau = tx.t0 + (sqrt_rn((ax * ax + az * az) + py2) + d);
                    tau = tau + a.offset;
                    const float v = interp_at<RFT, true>(rsrc, off, a.n_s, tau);
                    if (pass == 0) {
                        wsum += w;
                        nadd++;
                        sample += w * v;
                    } else {
                        acc.add(w, v, das);
                    }
                }
            }
            if (pass == 1) {
                const float s_local = acc.result(das, wsum, nadd);
                if (INTERP_TX) s_interp += sc.txw(k) * s_local;
                else s_interp += s_local;
            }
        }
    }
    if (valid) {
        OutT* out = static_cast<OutT*>(a.out) + static_cast<uint64_t>(frame) * a.out_frame_stride;
        out[static_cast<uint64_t>(r) * a.n_rx + s] = clamp_cast<OutT>(s_interp);
    }
}

template <typename RFT, typename OutT, int KIND>
void launch_kind(const DasArgs& a, bool is3d, bool interp_tx, dim3 grid, hipStream_t stream) {
    if (is3d) {
        if (interp_tx) hipLaunchKernelGGL((das3d_variant_kernel<RFT, OutT, KIND, true>), grid, dim3(kThreads), 0, stream, a);
        else hipLaunchKernelGGL((das3d_variant_kernel<RFT, OutT, KIND, false>), grid, dim3(kThreads), 0, stream, a);
    } else {
        if (interp_tx) hipLaunchKernelGGL((das2d_variant_kernel<RFT, OutT, KIND, true>), grid, dim3(kThreads), 0, stream, a);
        else hipLaunchKernelGGL((das2d_variant_kernel<RFT, OutT, KIND, false>), grid, dim3(kThreads), 0, stream, a);
    }
}

template <typename RFT, typename OutT>
void launch_types(const DasArgs& a, bool is3d, int kind, bool interp_tx, dim3 grid, hipStream_t stream) {
    if (kind == SUPRA_BF_DELAY_AND_STD_DEV) launch_kind<RFT, OutT, SUPRA_BF_DELAY_AND_STD_DEV>(a, is3d, interp_tx, grid, stream);
    else launch_kind<RFT, OutT, SUPRA_BF_COHERENCE_FACTOR>(a, is3d, interp_tx, grid, stream);
}

}  // namespace

hipError_t launch_das_variant(const DasArgs& a, int rf_dtype, int out_dtype, bool is3d, int kind, bool interp_tx,
                              hipStream_t stream) {
    // one workgroup per (frame, scanline, tile of kThreads depths); a.depth_tiles
    // is computed with the same tile (das_depth_tile() == kThreads)
    const uint64_t blocks = static_cast<uint64_t>(a.depth_tiles) * a.n_rx * a.n_frames;
    const dim3 grid(static_cast<uint32_t>(blocks));
    if (rf_dtype == SUPRA_INT16) {
        if (out_dtype == SUPRA_FLOAT32) launch_types<int16_t, float>(a, is3d, kind, interp_tx, grid, stream);
        else launch_types<int16_t, int16_t>(a, is3d, kind, interp_tx, grid, stream);
    } else {
        if (out_dtype == SUPRA_FLOAT32) launch_types<float, float>(a, is3d, kind, interp_tx, grid, stream);
        else launch_types<float, int16_t>(a, is3d, kind, interp_tx, grid, stream);
    }
    return hipGetLastError();
}

}  // namespace supra_amd
