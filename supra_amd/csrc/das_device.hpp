// Device helpers shared by the DAS forward and adjoint kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include <cfloat>
#include <cmath>
#include <cstddef>
#include <cstdint>

#include "bf_internal.hpp"

namespace supra_amd {
namespace {

constexpr int kBufferDword3 = 0x00020000;       // gfx9 raw buffer descriptor word 3

// clampCast (src/SupraLib/utilities/cudaUtility.h:111-147): fmaxf/fminf keep
// the non-NaN operand, so a NaN input lands on the lower limit.
template <typename T> __device__ inline T clamp_cast(float x);
template <> __device__ inline float clamp_cast<float>(float x) {
    return fminf(fmaxf(x, -FLT_MAX), FLT_MAX);
}
template <> __device__ inline int16_t clamp_cast<int16_t>(float x) {
    return static_cast<int16_t>(fminf(fmaxf(x, -32767.0f), 32767.0f));
}

// WindowFunctionGpu::get (WindowFunction.h:58-64): roundf of a value in
// [0, 63], computed as truncation plus a half-up step (identical to roundf
// for non-negative arguments).
__device__ inline int window_index(float rel) {
    const float x = 31.5f * (fminf(fmaxf(rel, -1.0f), 1.0f) + 1.0f);
    const uint32_t t = static_cast<uint32_t>(x);
    return static_cast<int>(t + ((x - static_cast<float>(t)) >= 0.5f ? 1u : 0u));
}

// Correctly rounded single-precision square root for x == 0 or normal x
// (every argument here is a sum of squares of dt-space lengths, so it is
// either exactly zero or far above 2^-96).  v_sqrt_f32 is within one ulp;
// one residual test on each neighbour picks the correctly rounded value
// (the correction LLVM emits for sqrtf, without the denormal rescale and the
// class test that this domain does not need).
__device__ inline float sqrt_rn(float x) {
    float s = __builtin_amdgcn_sqrtf(x);
    const uint32_t si = __float_as_uint(s);
    const float down = __uint_as_float(si - 1u);
    const float up = __uint_as_float(si + 1u);
    const float r_down = fmaf(-down, s, x);
    const float r_up = fmaf(-up, s, x);
    s = (r_down <= 0.0f) ? down : s;
    s = (r_up > 0.0f) ? up : s;
    return s;
}

// Workgroup-uniform tables are read through the constant address space so
// that they arrive in SGPRs (scalar loads); the buffer descriptor built from
// them is then uniform and the gathers need no per-lane descriptor loop.
typedef const __attribute__((address_space(4))) uint32_t* UniformWords;

// Field accessors of one ScanlineDT record through scalar loads (no copy of
// the record into registers or private memory; transmit k may be dynamic).
struct ScanView {
    UniformWords w;
    __device__ float f(size_t byte_offset) const { return __uint_as_float(w[byte_offset / 4]); }
    __device__ uint32_t u(size_t byte_offset) const { return w[byte_offset / 4]; }
    __device__ float x() const { return f(offsetof(ScanlineDT, x)); }
    __device__ float z() const { return f(offsetof(ScanlineDT, z)); }
    __device__ float dir_x() const { return f(offsetof(ScanlineDT, dir_x)); }
    __device__ float dir_y() const { return f(offsetof(ScanlineDT, dir_y)); }
    __device__ float dir_z() const { return f(offsetof(ScanlineDT, dir_z)); }
    __device__ float max_ed_x() const { return f(offsetof(ScanlineDT, max_ed_x)); }
    __device__ float max_ed_y() const { return f(offsetof(ScanlineDT, max_ed_y)); }
    __device__ int32_t sel() const { return static_cast<int32_t>(u(offsetof(ScanlineDT, sel))); }
    __device__ int32_t active_mask() const { return static_cast<int32_t>(u(offsetof(ScanlineDT, active_mask))); }
    __device__ float txw(int k) const { return __uint_as_float(w[offsetof(ScanlineDT, txw) / 4 + k]); }
    __device__ TxDT tx(int k) const {
        const size_t b = (offsetof(ScanlineDT, tx) + static_cast<size_t>(k) * sizeof(TxDT)) / 4;
        TxDT t;
        const uint32_t w0 = w[b], w1 = w[b + 1];
        t.first_x = static_cast<uint16_t>(w0 & 0xffffu);
        t.first_y = static_cast<uint16_t>(w0 >> 16);
        t.last_x = static_cast<uint16_t>(w1 & 0xffffu);
        t.last_y = static_cast<uint16_t>(w1 >> 16);
        t.tx_idx = w[b + 2];
        t.t0 = __uint_as_float(w[b + 3]);
        t.far_thr = __uint_as_float(w[b + 4]);
        return t;
    }
};

__device__ inline ScanView scan_view(const ScanlineDT* table, uint32_t s) {
    return ScanView{(UniformWords)(table + s)};
}

__device__ inline float load_uniform_float(const float* table, uint32_t i) {
    return __uint_as_float(((UniformWords)table)[i]);
}

// readfirstlane returns int: each half goes back through uint32_t before it
// is widened, or a low word with bit 31 set would sign-extend over the high word.
__device__ inline uint32_t uniform_u32(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
}

__device__ inline __amdgpu_buffer_rsrc_t rf_resource(const void* base, uint32_t bytes) {
    const uint64_t v = reinterpret_cast<uint64_t>(base);
    const uint64_t u = (static_cast<uint64_t>(uniform_u32(static_cast<uint32_t>(v >> 32))) << 32) |
                       static_cast<uint64_t>(uniform_u32(static_cast<uint32_t>(v)));
    return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void*>(u), 0, static_cast<int>(uniform_u32(bytes)),
                                             kBufferDword3);
}

template <typename T> __device__ inline float load_rf(__amdgpu_buffer_rsrc_t r, uint32_t idx);
template <> __device__ inline float load_rf<int16_t>(__amdgpu_buffer_rsrc_t r, uint32_t idx) {
    return static_cast<float>(static_cast<int16_t>(
        __builtin_amdgcn_raw_buffer_load_b16(r, static_cast<int>(idx * 2u), 0, 0)));
}
template <> __device__ inline float load_rf<float>(__amdgpu_buffer_rsrc_t r, uint32_t idx) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, static_cast<int>(idx * 4u), 0, 0));
}

// Linear interpolation of one channel at delay tau, with the reference's
// range rules (RxSampleBeamformerDelayAndSum.h:174-187 for 2D, :147-161 for
// 3D): i < N-1 interpolates, i == N-1 with zero fraction takes the sample,
// anything else contributes nothing.  SAT_UINT selects the 3D conversion of
// floor(tau) (saturating at 0, so negative delays read sample 0 with a
// negative fraction) versus the 2D int32 conversion (negative delays are out
// of range after the unsigned compare).
template <typename RFT, bool SAT_UINT>
__device__ inline void gather(__amdgpu_buffer_rsrc_t rf, uint32_t off, uint32_t n_s, float tau, float w,
                              float& sample) {
    uint32_t ui;
    float f;
    if constexpr (SAT_UINT) {
        // floor(tau) clamped at 0 (NaN -> 0 like the compare-and-select),
        // and tau minus that same float (equal to tau - float(ui) wherever
        // ui can be in range)
        const float fl = fmaxf(floorf(tau), 0.0f);
        ui = static_cast<uint32_t>(fl);
        f = tau - fl;
    } else {
        // tau - floor(tau) by v_fract_f32: exact for tau >= 0, the only case
        // that can be in range
        ui = static_cast<uint32_t>(static_cast<int32_t>(floorf(tau)));
        f = __builtin_amdgcn_fractf(tau);
    }
    if constexpr (std::is_same_v<RFT, int16_t>) {
        // int16 RF: "ui < N-1, or ui == N-1 with zero fraction" is a range
        // test on tau itself, and at tau == N-1 the neighbour sample is
        // finite, so (1 - 0) v0 + 0 v1 is v0 exactly and needs no separate
        // case.  2D (int32 floor): 0 <= tau <= N-1.  3D (floor saturated at
        // 0, so a negative tau interpolates from sample 0): tau <= N-1, or
        // tau == 0 when N == 1; a NaN delay, which the saturated floor maps
        // to sample 0, stays in range (for N > 1) as before.
        const float v0 = load_rf<RFT>(rf, off + ui);
        const float v1 = load_rf<RFT>(rf, off + ui + 1u);
        const float last = static_cast<float>(n_s - 1u);
        bool in;
        if constexpr (SAT_UINT) in = !(tau > last) & !(tau < (n_s > 1u ? -INFINITY : 0.0f));
        else in = (tau >= 0.0f) & (tau <= last);
        const float contrib = w * ((1.0f - f) * v0 + f * v1);
        sample += in ? contrib : 0.0f;   // (a select: keeps the loads unconditional, no branch per element)
        return;
    }
    const float v0 = load_rf<RFT>(rf, off + ui);
    const float v1 = load_rf<RFT>(rf, off + ui + 1u);
    const bool inr = ui < n_s - 1u;
    const bool edge = (ui == n_s - 1u) & (f == 0.0f);
    const float val = edge ? v0 : ((1.0f - f) * v0 + f * v1);
    const float contrib = w * val;
    sample = (inr | edge) ? sample + contrib : sample;
}

}  // namespace
}  // namespace supra_amd
