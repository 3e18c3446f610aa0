// Differentiable 2D delay-and-sum: per-frame dt-space tables built on the GPU
// for a device array of speeds of sound, and the adjoint of the beamformer
// with respect to the RF and to the speed of sound.
//
// No reference counterpart: SUPRA beamforms with one speed of sound per call
// and never differentiates (RxBeamformerCuda.cu:539-663).  The forward is the
// 2D kernel of das_kernels.hip (RxBeamformerCuda.cu:230-323 with
// RxSampleBeamformerDelayAndSum.h:122-208); for a fixed speed of sound its
// output is linear in the RF,
//     y(s,r) = sum_k a_k (n_k / W_k) sum_e w_e [(1-f) rf_k,ch(e)[i] + f rf_k,ch(e)[i+1]],
// with i = floor(tau), f = tau - i, a_k = txWeight_k (transmit interpolation)
// or 1 (the selected transmit).  The forward optionally writes
// norm = a_k n_k / W_k per pixel and slot; the adjoint scatters
// beta = dL/dy * norm times w_e (1-f) and w_e f back into the two samples.
//
// Speed of sound: every dt-space length is millimetres / (c dt), while the
// transmit delay t0 (seconds / dt) and the additional offset are fixed, so
// tau = t0 + G / c + offset and dtau/dc = -geo / c with geo = sqrt(q) + d,
// the geometric part the kernel evaluates anyway.  Aperture membership, the
// window argument and n/W are ratios of dt-space lengths: independent of c up
// to rounding, piecewise constant, derivative 0.  Hence
//     dL/dc = -(1/c) sum beta w_e (rf[i+1] - rf[i]) geo        (interpolated samples)
// and the delay tables are recomputed here, never stored by the forward.
//
// Launch layout (MI355X): one wave64 per (frame, scanline, slot, aperture
// element), 4 waves per workgroup.  A wave walks all depths of its element,
// lanes = 64 consecutive depths, and accumulates the element's whole RF line
// (num_samples floats) in a private LDS buffer with LDS float atomics; the
// delays are monotonic in depth, so the lanes of one instruction touch ~64
// neighbouring samples.  The finished line is written to HBM once, with
// plain coalesced stores when no other wave writes that (transmit, channel)
// line (each focused transmit feeds one scanline), else with global float
// atomics.  d/dc partial sums go to a per-wave slot and are reduced per frame
// in a fixed order, so the result does not depend on scheduling.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstddef>
#include <cstdint>

#include "bf_internal.hpp"
#include "das_device.hpp"
#include "das_launch.hpp"

namespace supra_amd {

namespace {

constexpr int kTableThreads = 256;
constexpr int kBetaTile = 64;
constexpr int kWave = 64;

// ---- per-frame tables --------------------------------------------------------

// The same double-precision expressions and float casts as build_dt_tables
// (bf_context.hip), which restates RxBeamformerCuda::convertToDtSpace
// (RxBeamformerCuda.cu:58-107) with the first-call factors.
__global__ __launch_bounds__(kTableThreads) void frame_tables_kernel(const FrameTablesArgs a) {
    const uint32_t f = blockIdx.y;
    const uint32_t i = blockIdx.x * kTableThreads + threadIdx.x;
    const double c = a.c[f];
    const double factor = 1 / 1.0 / (c * a.dt);
    const double factor_time = 1 / 1.0 / a.dt;
    if (i == 0) a.frame_c[f] = static_cast<float>(c);
    if (i < a.n_d) {
        a.depths[static_cast<size_t>(f) * a.n_d + i] = static_cast<float>(a.depths_mm[i] * factor);
        return;
    }
    uint32_t j = i - a.n_d;
    if (j < a.n_el) {
        a.ex[static_cast<size_t>(f) * a.n_el + j] = static_cast<float>(a.ex_mm[j] * factor);
        return;
    }
    j -= a.n_el;
    if (j >= a.n_rx) return;
    const supra_scanline_rx& s = a.scan_mm[j];
    ScanlineDT o = a.scan_fixed[j];
    o.x = static_cast<float>(s.position[0] * factor);
    o.z = static_cast<float>(s.position[2] * factor);
    o.max_ed_x = static_cast<float>(s.max_element_distance[0] * factor);
    o.max_ed_y = static_cast<float>(s.max_element_distance[1] * factor);
    for (int k = 0; k < 4; k++) {
        TxDT& t = o.tx[k];
        t.t0 = static_cast<float>(s.tx[k].initial_delay * factor_time);
        float thr = 0.0f;
        for (uint32_t e = t.first_x; e < t.last_x; e++) {
            if (a.e2c && a.e2c[e] < 0) continue;
            const float v = fabsf(static_cast<float>(a.ex_mm[e] * factor) - o.x);
            thr = (thr < v) ? v : thr;
        }
        t.far_thr = (thr < o.max_ed_x) ? o.max_ed_x : thr;
    }
    a.scans[static_cast<size_t>(f) * a.n_rx + j] = o;
}

// ---- beta = dL/dy * norm, transposed to [frame][slot][scanline][depth] -------

__global__ __launch_bounds__(256) void beta_kernel(const float* grad_image, const float* norm, float* beta,
                                                   uint32_t n_rx, uint32_t n_d, uint32_t slots) {
    __shared__ float tile[kBetaTile][kBetaTile + 1];
    const uint32_t s0 = blockIdx.x * kBetaTile, r0 = blockIdx.y * kBetaTile, f = blockIdx.z;
    const uint32_t tx = threadIdx.x % kBetaTile, ty = threadIdx.x / kBetaTile;
    const float* g = grad_image + static_cast<size_t>(f) * n_d * n_rx;
    for (uint32_t rr = ty; rr < kBetaTile; rr += 256 / kBetaTile) {
        const uint32_t r = r0 + rr, s = s0 + tx;
        tile[rr][tx] = (r < n_d && s < n_rx) ? g[static_cast<size_t>(r) * n_rx + s] : 0.0f;
    }
    __syncthreads();
    for (uint32_t q = 0; q < slots; q++) {
        const size_t base = (static_cast<size_t>(f) * slots + q) * n_rx;
        for (uint32_t ss = ty; ss < kBetaTile; ss += 256 / kBetaTile) {
            const uint32_t s = s0 + ss, r = r0 + tx;
            if (s < n_rx && r < n_d) {
                const size_t idx = (base + s) * n_d + r;
                const float nv = norm[idx];
                // Saturated pixels (norm 0) pass no gradient, whatever dL/dy is.
                beta[idx] = (nv == 0.0f) ? 0.0f : tile[tx][ss] * nv;
            }
        }
    }
}

// ---- the adjoint ---------------------------------------------------------------

__device__ inline double wave_sum(double v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

__global__ __launch_bounds__(256) void das2d_adjoint_kernel(const AdjointArgs a) {
    extern __shared__ float s_mem[];
    float* s_lut = s_mem;
    const uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    float* s_line = s_mem + kWindowEntries + static_cast<size_t>(wave) * a.n_s;
    for (uint32_t i = threadIdx.x; i < kWindowEntries; i += blockDim.x) s_lut[i] = a.lut[i];

    uint32_t b = blockIdx.x;
    const uint32_t eg = b % a.egroups;
    b /= a.egroups;
    const uint32_t q = b % a.n_slots;
    b /= a.n_slots;
    const uint32_t s = b % a.n_rx;
    const uint32_t f = b / a.n_rx;

    const ScanView sc = scan_view(a.scans + static_cast<size_t>(f) * a.scan_frame_stride, s);
    int k;
    bool active;
    if (a.interp_tx) {
        k = static_cast<int>(q);
        active = ((sc.active_mask() >> k) & 1) != 0;
    } else {
        k = sc.sel();
        active = k >= 0;
    }
    TxDT tx{};
    if (active) tx = sc.tx(k);
    const int e = static_cast<int>(tx.first_x) + static_cast<int>(eg * a.waves + wave);
    active = active && e < static_cast<int>(tx.last_x);
    uint32_t ch = 0;
    if (active) {
        if (a.e2c) {
            const int32_t m = a.e2c[e];
            active = m >= 0;
            ch = active ? static_cast<uint32_t>(m) : 0u;
        } else {
            ch = static_cast<uint32_t>(e) % a.n_ch;
        }
    }
    for (uint32_t i = lane; i < a.n_s; i += kWave) s_line[i] = 0.0f;
    __syncthreads();

    const size_t line = a.rf_frame_stride * f + (static_cast<size_t>(tx.tx_idx) * a.n_ch + ch) * a.n_s;
    float dc = 0.0f;
    if (active) {
        const float* depths = a.depths + static_cast<size_t>(f) * a.depth_frame_stride;
        const float* beta = a.beta + ((static_cast<size_t>(f) * a.n_slots + q) * a.n_rx + s) * a.n_depths;
        const float c = a.frame_c[f];
        const float x = sc.x(), dir_x = sc.dir_x(), dir_y = sc.dir_y(), med = sc.max_ed_x();
        const float xe = a.ex[static_cast<size_t>(f) * a.el_frame_stride + e];
        const float dx = xe - x;
        const float* rf_line = a.rf + line;
        const bool want_dc = a.dc_part != nullptr;
        // Lanes are consecutive depths, so their delays (never decreasing with
        // depth) are sorted across the wave.  A lane whose upper sample ui+1
        // is the next lane's lower sample hands its share to that lane, which
        // adds it before its own LDS add: one LDS float atomic instead of two
        // for that sample (the atomics' throughput bounds this kernel).
        for (uint32_t r0 = 0; r0 < a.n_depths; r0 += kWave) {
            const uint32_t r = r0 + lane;
            const bool valid = r < a.n_depths;
            const float bt = valid ? beta[r] : 0.0f;
            const float d = valid ? depths[r] : 0.0f;
            // Same single-precision operations as the forward kernel.
            const float aDT = (((d * a.dt) * c) / (2.0f * a.F)) / c / a.dt;
            const bool take = valid && bt != 0.0f && fabsf(dx) <= aDT;
            const float inv = 1.0f / fminf(aDT, med);
            const float w = s_lut[window_index(dx * inv)];
            const float px = x + dir_x * d;
            const float py = dir_y * d;
            const float ee = xe - px;
            const float geo = sqrt_rn(ee * ee + py * py) + d;
            float tau = tx.t0 + geo;
            tau = tau + a.offset;
            const int32_t i = static_cast<int32_t>(floorf(tau));
            const float fr = tau - static_cast<float>(i);
            const uint32_t ui = static_cast<uint32_t>(i);
            const float bw = bt * w;
            const bool inr = take && ui < a.n_s - 1u;
            const bool edge = take && ui == a.n_s - 1u && fr == 0.0f;
            const float lo = inr ? bw * (1.0f - fr) : (edge ? bw : 0.0f);
            const float hi = inr ? bw * fr : 0.0f;
            // link(L -> L+1) = inr(L) && valid(L+1) && ui(L+1) == ui(L) + 1, evaluated
            // identically by both lanes
            const uint32_t next_ui = __shfl_down(ui, 1, kWave);
            const bool next_valid = __shfl_down(valid ? 1 : 0, 1, kWave) != 0 && lane + 1 < kWave;
            const bool to_next = inr && next_valid && next_ui == ui + 1u;
            const uint32_t prev_ui = __shfl_up(ui, 1, kWave);
            const bool prev_inr = __shfl_up(inr ? 1 : 0, 1, kWave) != 0 && lane > 0;
            const float prev_hi = __shfl_up(hi, 1, kWave);
            const bool from_prev = prev_inr && valid && ui == prev_ui + 1u;
            if (inr || edge || from_prev) atomicAdd(&s_line[ui], from_prev ? lo + prev_hi : lo);
            if (inr && !to_next) atomicAdd(&s_line[ui + 1u], hi);
            if (want_dc && inr) dc += (bw * (rf_line[ui + 1u] - rf_line[ui])) * geo;
        }
    }
    __syncthreads();
    if (active && a.grad_rf) {
        float* out = a.grad_rf + line;
        if (a.exclusive) {
            for (uint32_t i = lane; i < a.n_s; i += kWave) out[i] = s_line[i];
        } else {
            for (uint32_t i = lane; i < a.n_s; i += kWave) {
                const float v = s_line[i];
                if (v != 0.0f) atomicAdd(&out[i], v);
            }
        }
    }
    if (a.dc_part) {
        const double v = wave_sum(static_cast<double>(dc));
        if (lane == 0) a.dc_part[static_cast<size_t>(blockIdx.x) * a.waves + wave] = v;
    }
}

// dL/dc per frame: fixed-order sum of the waves' partial sums, times -1/c.
__global__ __launch_bounds__(256) void dc_reduce_kernel(const double* part, uint32_t per_frame, const double* c,
                                                        double* grad_c) {
    __shared__ double red[256];
    const uint32_t f = blockIdx.x;
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < per_frame; i += 256) acc += part[static_cast<size_t>(f) * per_frame + i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) grad_c[f] = -red[0] / c[f];
}

// Zero the listed (transmit, channel) lines of every frame (lines no wave writes).
__global__ __launch_bounds__(256) void zero_lines_kernel(float* grad_rf, const uint32_t* lines, uint32_t n_s,
                                                         uint64_t frame_stride) {
    float* p = grad_rf + frame_stride * blockIdx.y + static_cast<size_t>(lines[blockIdx.x]) * n_s;
    for (uint32_t i = threadIdx.x; i < n_s; i += 256) p[i] = 0.0f;
}

__global__ __launch_bounds__(256) void fill_zero_kernel(float* p, uint64_t n) {
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * 256)
        p[i] = 0.0f;
}

}  // namespace

hipError_t launch_frame_tables(const FrameTablesArgs& a, hipStream_t stream) {
    const uint32_t items = a.n_d + a.n_el + a.n_rx;
    dim3 grid((items + kTableThreads - 1) / kTableThreads, a.n_frames), block(kTableThreads);
    hipLaunchKernelGGL(frame_tables_kernel, grid, block, 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_beta(const float* grad_image, const float* norm, float* beta, uint32_t n_rx, uint32_t n_d,
                       uint32_t slots, uint32_t frames, hipStream_t stream) {
    dim3 grid((n_rx + kBetaTile - 1) / kBetaTile, (n_d + kBetaTile - 1) / kBetaTile, frames), block(256);
    hipLaunchKernelGGL(beta_kernel, grid, block, 0, stream, grad_image, norm, beta, n_rx, n_d, slots);
    return hipGetLastError();
}

size_t das_adjoint_lds_bytes(uint32_t waves, uint32_t n_s) {
    return (static_cast<size_t>(kWindowEntries) + static_cast<size_t>(waves) * n_s) * sizeof(float);
}

hipError_t launch_das_adjoint(const AdjointArgs& a, uint32_t frames, hipStream_t stream) {
    const uint64_t blocks = static_cast<uint64_t>(frames) * a.n_rx * a.n_slots * a.egroups;
    dim3 grid(static_cast<uint32_t>(blocks)), block(a.waves * kWave);
    hipLaunchKernelGGL(das2d_adjoint_kernel, grid, block, das_adjoint_lds_bytes(a.waves, a.n_s), stream, a);
    return hipGetLastError();
}

hipError_t launch_dc_reduce(const double* part, uint32_t per_frame, const double* c, double* grad_c, uint32_t frames,
                            hipStream_t stream) {
    hipLaunchKernelGGL(dc_reduce_kernel, dim3(frames), dim3(256), 0, stream, part, per_frame, c, grad_c);
    return hipGetLastError();
}

hipError_t launch_zero_lines(float* grad_rf, const uint32_t* lines, uint32_t n_lines, uint32_t n_s,
                             uint64_t frame_stride, uint32_t frames, hipStream_t stream) {
    if (n_lines == 0) return hipSuccess;
    hipLaunchKernelGGL(zero_lines_kernel, dim3(n_lines, frames), dim3(256), 0, stream, grad_rf, lines, n_s,
                       frame_stride);
    return hipGetLastError();
}

hipError_t launch_fill_zero(float* p, uint64_t n, hipStream_t stream) {
    const uint64_t want = (n + 255) / 256;
    const uint32_t blocks = static_cast<uint32_t>(want < 8192 ? (want ? want : 1) : 8192);
    hipLaunchKernelGGL(fill_zero_kernel, dim3(blocks), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

}  // namespace supra_amd
