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
// lanes = 64 consecutive depths, and accumulates the element's RF line with
// LDS float atomics in a ring of 256 samples: the delays grow with depth, so
// the lanes of one instruction touch ~64 neighbouring samples, and samples
// more than a guard below the delay of the chunk's first depth are final.
// Those leave the ring in order, as coalesced stores when no other wave
// writes that (transmit, channel) line (each focused transmit feeds one
// scanline), else as global float atomics; a target outside the ring's window
// (delays not sorted, or a chunk wider than the ring) is handled exactly by
// moving the ring or adding to the flushed output.  1 KB of LDS per wave
// instead of a whole line leaves room for the maximum number of waves per
// SIMD.  d/dc partial sums go to a per-wave slot and are reduced per frame
// in a fixed order, so the result does not depend on scheduling.  The
// depth-only quantities (the F-number aperture aDT and 1/aDT, four IEEE
// divisions per depth) come from a per-frame table built once, not once per
// aperture element; neighbouring lanes exchange their sample indices by DPP
// moves, which leave the LDS pipe to the atomics.
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
constexpr uint32_t kThrParts = 4;                // element strides per transmit slot
constexpr uint32_t kScanLanes = 4 * kThrParts;   // lanes per scanline in frame_tables_kernel

__host__ __device__ inline uint32_t scan_items_start(uint32_t n_d, uint32_t n_el) {
    return (n_d + n_el + 63u) / 64u * 64u;
}
constexpr int kBetaTile = 32;
constexpr int kWave = 64;
// Each adjoint wave accumulates its RF line in a ring of kRing samples of LDS;
// samples more than kRingGuard below the delay of the chunk's first depth are
// final and leave the ring.
constexpr uint32_t kRing = 256;
constexpr int32_t kRingGuard = 32;

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
    // Scanlines: kScanLanes consecutive lanes per scanline (4 transmit slots x
    // kThrParts element strides), from the first item at a multiple of the
    // wave size, so that each group is one aligned shuffle segment; the
    // far-field threshold of a slot is the largest |ex - x| over its
    // connected elements, a max, so the parts reduce in any order.
    const uint32_t s0 = scan_items_start(a.n_d, a.n_el);
    if (i < s0) return;
    j = i - s0;
    const uint32_t sub = j % kScanLanes;
    j /= kScanLanes;
    if (j >= a.n_rx) return;   // whole groups of kScanLanes lanes
    const supra_scanline_rx& s = a.scan_mm[j];
    ScanlineDT o = a.scan_fixed[j];
    o.x = static_cast<float>(s.position[0] * factor);
    o.z = static_cast<float>(s.position[2] * factor);
    o.max_ed_x = static_cast<float>(s.max_element_distance[0] * factor);
    o.max_ed_y = static_cast<float>(s.max_element_distance[1] * factor);
    const uint32_t kk = sub / kThrParts, part = sub % kThrParts;
    float thr = 0.0f;
    {
        const TxDT& t = o.tx[kk];
        for (uint32_t e = t.first_x + part; e < t.last_x; e += kThrParts) {
            if (a.e2c && a.e2c[e] < 0) continue;
            const float v = fabsf(static_cast<float>(a.ex_mm[e] * factor) - o.x);
            thr = (thr < v) ? v : thr;
        }
    }
    for (uint32_t m = 1; m < kThrParts; m <<= 1) {
        const float v = __shfl_xor(thr, static_cast<int>(m), kScanLanes);
        thr = (thr < v) ? v : thr;
    }
    float thr_k[4];
    for (int k = 0; k < 4; k++) thr_k[k] = __shfl(thr, k * static_cast<int>(kThrParts), kScanLanes);
    if (sub != 0) return;
    for (int k = 0; k < 4; k++) {
        TxDT& t = o.tx[k];
        t.t0 = static_cast<float>(s.tx[k].initial_delay * factor_time);
        t.far_thr = (thr_k[k] < o.max_ed_x) ? o.max_ed_x : thr_k[k];
    }
    a.scans[static_cast<size_t>(f) * a.n_rx + j] = o;
}

// ---- per-frame depth table ----------------------------------------------------

// {depth, aDT, 1/aDT, 0} per (frame, depth), aDT with the same single-precision
// operations as the forward kernel (das_kernels.hip, after
// RxSampleBeamformerDelayAndSum.h:167 and RxBeamformerCommon.h:27-33).
__global__ __launch_bounds__(256) void depth_table_kernel(const float* depths, const float* frame_c, float dt,
                                                          float f_number, uint32_t n_d, float* dtab) {
    const uint32_t f = blockIdx.y;
    const uint32_t r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n_d) return;
    const size_t i = static_cast<size_t>(f) * n_d + r;
    const float d = depths[i];
    const float c = frame_c[f];
    const float aDT = (((d * dt) * c) / (2.0f * f_number)) / c / dt;
    reinterpret_cast<float4*>(dtab)[i] = make_float4(d, aDT, 1.0f / aDT, 0.0f);
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

// Whole-wave DPP shifts by one lane (wave_shl:1, wave_shr:1): lane i reads
// lane i + 1 (from_next) or lane i - 1 (from_prev); lane 63, resp. lane 0,
// reads 0.  Called with every lane of the wave active.
__device__ inline uint32_t from_next(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x130, 0xf, 0xf, false));
}
__device__ inline uint32_t from_prev(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x138, 0xf, 0xf, false));
}

// Largest wave max of v over the wave (rare paths only).
__device__ inline uint32_t wave_max(uint32_t v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v = max(v, static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), o, kWave)));
    return v;
}

// Moves samples [from, to) of a wave's RF line from its ring to the output
// line, in order, and clears their ring slots.  Samples at or beyond
// from + kRing were never added to (every target lies below base + kRing when
// it is added and base only grows), so they are written as 0.
__device__ inline void ring_flush(float* ring, float* out, bool exclusive, uint32_t lane, int32_t from, int32_t to) {
    for (int32_t u = from + static_cast<int32_t>(lane); u < to; u += kWave) {
        float v = 0.0f;
        if (u < from + static_cast<int32_t>(kRing)) {
            float& slot = ring[static_cast<uint32_t>(u) & (kRing - 1u)];
            v = slot;
            slot = 0.0f;
        }
        if (exclusive) out[u] = v;
        else if (v != 0.0f) atomicAdd(&out[u], v);
    }
}

__global__ __launch_bounds__(256) void das2d_adjoint_kernel(const AdjointArgs a) {
    extern __shared__ float s_mem[];
    float* s_lut = s_mem;
    const uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    float* s_ring = s_mem + kWindowEntries + static_cast<size_t>(wave) * kRing;
    for (uint32_t i = threadIdx.x; i < kWindowEntries; i += blockDim.x) s_lut[i] = a.lut[i];
    for (uint32_t i = lane; i < kRing; i += kWave) s_ring[i] = 0.0f;

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
    __syncthreads();

    const size_t line = a.rf_frame_stride * f + (static_cast<size_t>(tx.tx_idx) * a.n_ch + ch) * a.n_s;
    float dc = 0.0f;
    if (active && a.n_depths > 0u) {
        const float4* dtab =
            reinterpret_cast<const float4*>(a.dtab) + static_cast<size_t>(f) * a.depth_frame_stride;
        const float* beta = a.beta + ((static_cast<size_t>(f) * a.n_slots + q) * a.n_rx + s) * a.n_depths;
        const float x = sc.x(), dir_x = sc.dir_x(), dir_y = sc.dir_y(), med = sc.max_ed_x();
        // 1 / fminf(aDT, med) is 1/aDT below med and 1/med from there on.
        const float inv_med = 1.0f / med;
        const float xe = a.ex[static_cast<size_t>(f) * a.el_frame_stride + e];
        const float dx = xe - x;
        const float* rf_line = a.rf + line;
        const bool want_dc = a.dc_part != nullptr;
        float* out = a.grad_rf ? a.grad_rf + line : nullptr;
        const bool exclusive = a.exclusive != 0;
        const int32_t n_s = static_cast<int32_t>(a.n_s);
        // Samples below base are final and in out[]; [base, base + kRing)
        // are accumulated in the ring (slot = sample mod kRing).
        int32_t base = 0;
        const uint32_t last = a.n_depths - 1u;
        // This iteration's depth entry and beta; the next iteration's are
        // loaded at the top of each iteration (clamped to the last depth).
        float4 dv = dtab[min(lane, last)];
        float btv = beta[min(lane, last)];
        // Lanes are consecutive depths, so their delays (never decreasing with
        // depth) are sorted across the wave.  A lane whose upper sample ui+1
        // is the next lane's lower sample hands its share to that lane, which
        // adds it with its own: one add instead of two for that sample.
        for (uint32_t r0 = 0; r0 < a.n_depths; r0 += kWave) {
            const uint32_t r = r0 + lane;
            const bool valid = r < a.n_depths;
            const float d = dv.x, aDT = dv.y;
            const float inv = aDT < med ? dv.z : inv_med;
            const float bt = valid ? btv : 0.0f;
            {
                const uint32_t rn = min(r + kWave, last);
                dv = dtab[rn];
                btv = beta[rn];
            }
            // Same single-precision operations as the forward kernel.
            const bool take = valid && bt != 0.0f && fabsf(dx) <= aDT;
            // No lane of this chunk in the element's aperture (shallow depths
            // of the outer elements): nothing to add; the ring catches up at
            // the next chunk that has work.
            if (__ballot(take) == 0ull) continue;
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
            float g0 = 0.0f, g1 = 0.0f;
            if (want_dc) {
                // the two samples of d/dc, issued before the LDS adds
                g0 = rf_line[inr ? ui : 0u];
                g1 = rf_line[inr ? ui + 1u : 0u];
            }
            if (out) {
                const float lo = inr ? bw * (1.0f - fr) : (edge ? bw : 0.0f);
                const float hi = inr ? bw * fr : 0.0f;
                // link(L -> L+1) = inr(L) && valid(L+1) && ui(L+1) == ui(L) + 1, evaluated
                // identically by both lanes
                const uint32_t next_ui = from_next(ui);
                const uint32_t next_fl = from_next((valid ? 1u : 0u) | (inr ? 2u : 0u));
                const bool next_valid = (next_fl & 1u) != 0u, next_inr = (next_fl & 2u) != 0u;
                const bool to_next = inr && next_valid && next_ui == ui + 1u;
                const uint32_t prev_ui = from_prev(ui);
                const bool prev_inr = from_prev(inr ? 1u : 0u) != 0u;
                const float prev_hi = __uint_as_float(from_prev(__float_as_uint(hi)));
                const bool from_prev_lane = prev_inr && valid && ui == prev_ui + 1u;
                const float lo_sum = from_prev_lane ? lo + prev_hi : lo;
                // Targets: ui (lo) and ui + 1 (hi), both below n_s.
                const bool add_lo = inr || edge || from_prev_lane;
                const bool add_hi = inr && !to_next;
                // Samples more than a guard below lane 0's (this chunk's
                // smallest delay) are final: flush them.
                const int32_t i0 = max(__builtin_amdgcn_readfirstlane(i), -1);
                const int32_t want = min(i0 - kRingGuard, n_s);
                if (want > base) {
                    ring_flush(s_ring, out, exclusive, lane, base, want);
                    base = want;
                }
                // A target beyond the ring moves the ring up to it (wide
                // apertures per chunk only).
                const uint32_t top = add_hi ? ui + 1u : (add_lo ? ui : 0u);
                if (__ballot(static_cast<int32_t>(top) >= base + static_cast<int32_t>(kRing)) != 0ull) {
                    const int32_t up = static_cast<int32_t>(__builtin_amdgcn_readfirstlane(wave_max(top))) -
                                       static_cast<int32_t>(kRing) + 1;
                    ring_flush(s_ring, out, exclusive, lane, base, up);
                    base = up;
                }
                const bool lo_ring = add_lo && static_cast<int32_t>(ui) >= base;
                const bool hi_ring = add_hi && static_cast<int32_t>(ui + 1u) >= base;
                const bool below = (add_lo && !lo_ring) || (add_hi && !hi_ring);
                // Atomic-free form.  A lane whose next lane has the same
                // sample (both in range) gives that lane its two shares, so
                // that every sample is written by at most one lane of each of
                // the two adds below: plain LDS read-add-write.  It holds
                // when the samples are sorted across the wave, no three
                // consecutive lanes share one and no other pair does.
                const bool same_next = inr && next_inr && next_ui == ui;
                const bool same_prev = prev_inr && inr && ui == prev_ui;
                const float a_own = lo_sum, b_own = add_hi ? hi : 0.0f;
                const float a_prev = __uint_as_float(from_prev(__float_as_uint(a_own)));
                const float a_tot = same_prev ? a_own + a_prev : a_own;
                const float b_tot = same_prev ? b_own + prev_hi : b_own;
                const bool do_a = add_lo && !same_next;
                const bool do_b = (add_hi && !same_next) || same_prev;
                const uint32_t prev_do = from_prev((do_a ? 1u : 0u) | (do_b ? 2u : 0u));
                const bool eq_prev = lane > 0u && valid && ui == prev_ui;
                const bool eq_next = valid && next_valid && next_ui == ui;
                const bool unsorted = lane > 0u && valid && i < static_cast<int32_t>(prev_ui);
                const bool dup = eq_prev && (((prev_do & 1u) != 0u && do_a) || ((prev_do & 2u) != 0u && do_b));
                if (__ballot(below || unsorted || (eq_prev && eq_next) || dup) == 0ull) {
                    if (do_a) s_ring[ui & (kRing - 1u)] += a_tot;
                    if (do_b) s_ring[(ui + 1u) & (kRing - 1u)] += b_tot;
                } else {
                    // A target below the ring (delays not sorted) is added to
                    // the flushed output directly.
                    if (__ballot(below) != 0ull) {
                        __threadfence();
                        if (add_lo && !lo_ring) atomicAdd(&out[ui], lo_sum);
                        if (add_hi && !hi_ring) atomicAdd(&out[ui + 1u], hi);
                    }
                    if (lo_ring) atomicAdd(&s_ring[ui & (kRing - 1u)], lo_sum);
                    if (hi_ring) atomicAdd(&s_ring[(ui + 1u) & (kRing - 1u)], hi);
                }
            }
            if (want_dc && inr) dc += (bw * (g1 - g0)) * geo;
        }
        if (out) ring_flush(s_ring, out, exclusive, lane, base, n_s);
    }
    if (a.dc_part) {
        const double v = wave_sum(static_cast<double>(dc));
        if (lane == 0) a.dc_part[static_cast<size_t>(blockIdx.x) * a.waves + wave] = v;
    }
}

// dL/dc per frame: fixed-order sum of the waves' partial sums, times -1/c.
constexpr uint32_t kReduceThreads = 1024;

__global__ __launch_bounds__(kReduceThreads) void dc_reduce_kernel(const double* part, uint32_t per_frame,
                                                                   const double* c, double* grad_c) {
    __shared__ double red[kReduceThreads];
    const uint32_t f = blockIdx.x;
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < per_frame; i += kReduceThreads)
        acc += part[static_cast<size_t>(f) * per_frame + i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t w = kReduceThreads / 2; w > 0; w >>= 1) {
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
    const uint32_t items = scan_items_start(a.n_d, a.n_el) + a.n_rx * kScanLanes;
    dim3 grid((items + kTableThreads - 1) / kTableThreads, a.n_frames), block(kTableThreads);
    hipLaunchKernelGGL(frame_tables_kernel, grid, block, 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_depth_table(const float* depths, const float* frame_c, float dt, float f_number, uint32_t n_d,
                              uint32_t frames, float* dtab, hipStream_t stream) {
    dim3 grid((n_d + 255) / 256, frames), block(256);
    hipLaunchKernelGGL(depth_table_kernel, grid, block, 0, stream, depths, frame_c, dt, f_number, n_d, dtab);
    return hipGetLastError();
}

hipError_t launch_beta(const float* grad_image, const float* norm, float* beta, uint32_t n_rx, uint32_t n_d,
                       uint32_t slots, uint32_t frames, hipStream_t stream) {
    dim3 grid((n_rx + kBetaTile - 1) / kBetaTile, (n_d + kBetaTile - 1) / kBetaTile, frames), block(256);
    hipLaunchKernelGGL(beta_kernel, grid, block, 0, stream, grad_image, norm, beta, n_rx, n_d, slots);
    return hipGetLastError();
}

size_t das_adjoint_lds_bytes(uint32_t waves) {
    return (static_cast<size_t>(kWindowEntries) + static_cast<size_t>(waves) * kRing) * sizeof(float);
}

hipError_t launch_das_adjoint(const AdjointArgs& a, uint32_t frames, hipStream_t stream) {
    const uint64_t blocks = static_cast<uint64_t>(frames) * a.n_rx * a.n_slots * a.egroups;
    dim3 grid(static_cast<uint32_t>(blocks)), block(a.waves * kWave);
    hipLaunchKernelGGL(das2d_adjoint_kernel, grid, block, das_adjoint_lds_bytes(a.waves), stream, a);
    return hipGetLastError();
}

hipError_t launch_dc_reduce(const double* part, uint32_t per_frame, const double* c, double* grad_c, uint32_t frames,
                            hipStream_t stream) {
    hipLaunchKernelGGL(dc_reduce_kernel, dim3(frames), dim3(kReduceThreads), 0, stream, part, per_frame, c, grad_c);
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
