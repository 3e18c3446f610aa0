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
// Launch layout (MI355X).  2D: each wave64 is an independent unit of 64
// consecutive depths of one (frame, rx scanline), with its own LDS tables
// and no workgroup barrier (a 2080-deep image is 32.5 units, so almost no
// lane idles in the last one).  3D: one workgroup of 256 lanes (4 wave64)
// per (frame, rx scanline, tile of 256 depths), aperture elements staged in
// LDS 128 at a time.  Lanes map to consecutive depths, so the two RF gathers
// of one element hit a contiguous window of the channel.  Per-element
// quantities that do not depend on depth are computed once per unit into
// LDS and broadcast.
//
// The kernels are VALU-bound (a correctly rounded sqrt and a linear
// interpolation per element and pixel), so the work per element is trimmed:
//  * far field, decided per wave: once the F-number aperture aDT at the
//    wave's first depth covers every aperture element and exceeds
//    maxElementDistance (aDT grows monotonically with depth), membership is
//    always true and the window argument no longer depends on depth, so each
//    element's weight is looked up once when it is staged;
//  * RF gathers are raw buffer loads (32-bit offsets, hardware range check
//    returns 0 beyond the transmit's block), issued unconditionally so that
//    the out-of-range cases of the reference become selects, not branches,
//    and several elements' loads are in flight at once.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <type_traits>

#include "bf_internal.hpp"
#include "das_device.hpp"
#include "das_launch.hpp"

namespace supra_amd {

namespace {

constexpr int kThreads = 256;
constexpr int kChunk2D = 128;     // aperture elements staged per LDS pass (2D, per wave)
// 2D work unit: one wave = 64 consecutive depths of one scanline (a 2080-deep
// image is 32.5 such units; with 256-depth workgroup tiles it would be 8.1
// tiles, a tenth of the lanes idle in the last one).
constexpr int kWaveDepths = 64;
constexpr int kWavesPerGroup = kThreads / 64;
// Frames per wave of the multi-frame 2D kernel (frames sharing a geometry).
#ifndef SUPRA_FRAMES_PER_WAVE
#define SUPRA_FRAMES_PER_WAVE 8
#endif
constexpr int kFramesPerWave = SUPRA_FRAMES_PER_WAVE;
// Aperture elements staged per LDS pass (3D).  Small on purpose: a small
// LDS footprint per workgroup (20 KB with the window table) lets 8
// workgroups, the 8 waves per SIMD that the registers allow, share a CU.
constexpr int kChunk3D = 128;
// The 3D window weight sqrt(LUT[ix] * LUT[iz]) as a full 64 x 64 table,
// entry ix * 64 + iz: one add per element with ix * 64 kept per element row
// (16 KB; with the element chunk a 3D workgroup takes 20 KB of LDS, and 8 of
// them, 8 waves per SIMD, still fit a CU).
constexpr int kW2Entries = kWindowEntries * kWindowEntries;
constexpr uint32_t kNotConnected = 0xffffffffu;

struct Pixel2D {
    float d, aDT, inv, px, py2;
};

// One 2D element's gathers, issued before any of a batch is consumed
// (gather<RFT, false>'s loads and decisions, RxSampleBeamformerDelayAndSum.h:
// 174-187).  A skipped element (outside the aperture, unconnected channel)
// has member == false and weight 0.
template <typename RFT>
struct Gather2D {
    float v0, v1, f, w;
    bool add;    // int16: member and 0 <= tau <= N-1; float: member and (inr | edge)
    bool edge;   // float: the last sample with zero fraction
};

template <typename RFT>
__device__ inline Gather2D<RFT> gather2d_issue(__amdgpu_buffer_rsrc_t rf, uint32_t off, uint32_t n_s, float tau,
                                               float w, bool member) {
    Gather2D<RFT> g;
    const uint32_t ui = static_cast<uint32_t>(static_cast<int32_t>(floorf(tau)));
    g.f = __builtin_amdgcn_fractf(tau);
    g.w = w;
    g.v0 = load_rf<RFT>(rf, off + ui);
    g.v1 = load_rf<RFT>(rf, off + ui + 1u);
    if constexpr (std::is_same_v<RFT, int16_t>) {
        const float last = static_cast<float>(n_s - 1u);
        g.add = member & (tau >= 0.0f) & (tau <= last);
        g.edge = false;
    } else {
        const bool inr = ui < n_s - 1u;
        g.edge = (ui == n_s - 1u) & (g.f == 0.0f);
        g.add = member & (inr | g.edge);
    }
    return g;
}

// The element's contribution in gather()'s operations; a skipped element adds
// +0 (int16: the sum is never -0) or leaves the sum selected unchanged.
template <typename RFT>
__device__ inline void gather2d_consume(const Gather2D<RFT>& g, float& sample) {
    if constexpr (std::is_same_v<RFT, int16_t>) {
        const float contrib = g.w * ((1.0f - g.f) * g.v0 + g.f * g.v1);
        sample += g.add ? contrib : 0.0f;
    } else {
        const float val = g.edge ? g.v0 : ((1.0f - g.f) * g.v0 + g.f * g.v1);
        const float contrib = g.w * val;
        sample = g.add ? sample + contrib : sample;
    }
}

// Elements [0, cnt) of the current LDS chunk for one 2D pixel
// (RxSampleBeamformerDelayAndSum.h:150-199), in batches of kBatch2D whose
// delays and gathers all go out before the first is consumed (as in 3D);
// weights, counts and samples are added in element order.
constexpr int kBatch2D = 4;

template <typename RFT, bool FAR, bool VERT, bool MAP>
__device__ inline Gather2D<RFT> element2d_issue(__amdgpu_buffer_rsrc_t rf, uint32_t n_s, float t0, float offset,
                                                const Pixel2D& p, const float* s_xe, const float* s_dx,
                                                const float* s_dx2, const float* s_w, const uint32_t* s_off,
                                                const float* s_lut, int j, float& wsum, int& nadd) {
    const uint32_t off = s_off[j];
    bool member = !MAP || off != kNotConnected;
    float w;
    if (FAR) {
        w = member ? s_w[j] : 0.0f;
    } else {
        const float dx = s_dx[j];
        member = member && (fabsf(dx) <= p.aDT);
        w = member ? s_lut[window_index(dx * p.inv)] : 0.0f;
    }
    // a skipped element adds +0 to wsum (never -0: window weights are >= 0) and 0 to n
    wsum += w;
    nadd += member ? 1 : 0;
    float q;
    if (VERT) {
        q = s_dx2[j] + p.py2;
    } else {
        const float e = s_xe[j] - p.px;
        q = e * e + p.py2;
    }
    float tau = t0 + (sqrt_rn(q) + p.d);
    tau = tau + offset;
    return gather2d_issue<RFT>(rf, off, n_s, tau, w, member);
}

template <typename RFT, bool FAR, bool VERT, bool MAP>
__device__ inline void elements2d(__amdgpu_buffer_rsrc_t rf, uint32_t n_s, float t0, float offset, const Pixel2D& p,
                                  const float* s_xe, const float* s_dx, const float* s_dx2, const float* s_w,
                                  const uint32_t* s_off, const float* s_lut, int cnt, float& sample, float& wsum,
                                  int& nadd) {
    int j = 0;
    for (; j + kBatch2D <= cnt; j += kBatch2D) {
        if (!FAR) {
            // near field: a batch that no lane of the wave has in its
            // aperture adds nothing (+0 weights, no samples) and is skipped
            bool any = false;
#pragma unroll
            for (int u = 0; u < kBatch2D; u++)
                any |= (fabsf(s_dx[j + u]) <= p.aDT) && (!MAP || s_off[j + u] != kNotConnected);
            if (__builtin_amdgcn_ballot_w64(any) == 0) continue;   // uniform
        }
        Gather2D<RFT> g[kBatch2D];
#pragma unroll
        for (int u = 0; u < kBatch2D; u++)
            g[u] = element2d_issue<RFT, FAR, VERT, MAP>(rf, n_s, t0, offset, p, s_xe, s_dx, s_dx2, s_w, s_off, s_lut,
                                                        j + u, wsum, nadd);
#pragma unroll
        for (int u = 0; u < kBatch2D; u++) gather2d_consume<RFT>(g[u], sample);
    }
    for (; j < cnt; j++) {
        const Gather2D<RFT> g = element2d_issue<RFT, FAR, VERT, MAP>(rf, n_s, t0, offset, p, s_xe, s_dx, s_dx2, s_w,
                                                                     s_off, s_lut, j, wsum, nadd);
        gather2d_consume<RFT>(g, sample);
    }
}

// RxSampleBeamformerTestSignal::sampleBeamform2D (RxSampleBeamformerTestSignal.h:87-133):
// 1000 inside 2 mm cylinders every 6 mm at 30 mm depth, 0 elsewhere.
__device__ inline float test_signal_2d(float x, float dir_x, float dir_y, float d, float dt, float c) {
    const float px = ((x + dir_x * d) * dt) * c;
    const float py = ((dir_y * d) * dt) * c;
    float sample = 0.0f;
    for (int k = -3; k <= 3; k++) {
        const float ex = px - static_cast<float>(k) * 6.0f, ey = py - 30.0f, ez = 0.0f;
        if (sqrtf(ex * ex + ey * ey + ez * ez) <= 2.0f) sample = 1000.0f;
    }
    return sample;
}

struct Tile {
    uint32_t s, frame, r, r0;
    bool valid;
};

__device__ inline Tile decode_tile(const DasArgs& a, uint32_t bid = blockIdx.x) {
    Tile t;
    const uint32_t tile = bid % a.depth_tiles;
    const uint32_t rest = bid / a.depth_tiles;
    t.s = rest % a.n_rx;
    t.frame = rest / a.n_rx;
    t.r0 = tile * kThreads;
    t.r = t.r0 + threadIdx.x;
    t.valid = t.r < a.n_depths;
    return t;
}

// The hardware deals workgroup b to XCD b % 8 (round robin), and each XCD
// has its own L2.  2D workgroups are renumbered in tiles of kXcdTileWgs
// consecutive logical workgroups (32 units: 32 neighbouring scanlines of
// one 64-depth chunk, which write the same output cache lines), dealt to
// the XCDs in rounds of 8 tiles: round j gives tile 8j + (x + j) % 8 to XCD
// x.  Each tile runs whole on one XCD, so the partial lines of an output
// row merge in its L2 instead of reaching HBM as separate writes, and the
// rotation spreads shallow and deep, central and edge tiles over all XCDs
// (a contiguous range per XCD left half of them with the shallow chunks and
// 27 % less work than the others).  A bijection on [0, gridDim.x): the
// launch pads the grid to whole rounds, padded units are past the end.
constexpr uint32_t kXcds = 8;
constexpr uint32_t kXcdTileWgs = 8;
__device__ inline uint32_t logical_wg_2d(uint32_t tile_wgs = kXcdTileWgs) {
    const uint32_t b = blockIdx.x;
    const uint32_t xcd = b % kXcds, local = b / kXcds;
    const uint32_t round = local / tile_wgs, k = local % tile_wgs;
    return (round * kXcds + (xcd + round) % kXcds) * tile_wgs + k;
}

// Multi-frame 2D tiles: 32 neighbouring scanlines x kTileChunks adjacent
// 64-depth chunks of one frame group.  A scanline's neighbouring chunks read
// the same RF lines at their common boundary; in one tile they run on one
// XCD and share them in its L2 (32 linear_2d frames: 8.9 -> 8.5 us per frame
// with 2; 3, 4 and 6 time the same).
constexpr uint32_t kTileChunks = 2;
constexpr uint32_t kTileLines = 32;

// The 2D kernels' unit: wave w of (renumbered) workgroup b takes unit
// b * 4 + w of the (64-depth chunk, frame, scanline) list, scanline fastest:
// lanes are depths, so an output row (consecutive scanlines at one depth) is
// written by neighbouring units.  Chunks run from the deepest to the
// shallowest: deep units have the widest aperture and take longest, and
// dispatching them first leaves short units for the end of the grid.  Units
// past the end are invalid.
__device__ inline Tile decode_wave_2d(const DasArgs& a) {
    const uint32_t chunks = (a.n_depths + kWaveDepths - 1) / kWaveDepths;
    const uint32_t wave = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / 64)));
    const uint64_t unit = static_cast<uint64_t>(logical_wg_2d()) * kWavesPerGroup + wave;
    const uint64_t rest = unit / a.n_rx;
    Tile t;
    t.s = static_cast<uint32_t>(unit % a.n_rx);
    const uint64_t cr = rest / a.n_frames;            // chunk, counted from the deepest
    const bool in_list = cr < chunks;
    t.frame = static_cast<uint32_t>(rest % a.n_frames);
    t.r0 = (in_list ? chunks - 1u - static_cast<uint32_t>(cr) : 0u) * kWaveDepths;
    t.r = t.r0 + threadIdx.x % 64;
    t.valid = in_list && t.r < a.n_depths;
    return t;
}

__device__ inline void wave_sync_2d() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS working set of one wave's 2D unit.
struct Lds2D {
    float lut[kWindowEntries];
    float xe[kChunk2D];
    float dx[kChunk2D];
    float dx2[kChunk2D];
    float w[kChunk2D];
    uint32_t off[kChunk2D];
};

// One wave's 2D unit: 64 consecutive depths (t.r0 ...) of one scanline of one
// frame, with the wave's own LDS tables (no workgroup barrier).
template <typename RFT, typename OutT, bool TESTSIG, bool INTERP_TX, bool MAP>
__device__ inline void das2d_tile(const DasArgs& a, const float* lut, const Tile& t, Lds2D& L) {
    if (!__any(t.valid)) return;   // (wave-uniform) a unit past the end of the image or the list
    const int lane = static_cast<int>(threadIdx.x % 64);
    float* s_lut = L.lut;
    float* s_xe = L.xe;
    float* s_dx = L.dx;
    float* s_dx2 = L.dx2;
    float* s_w = L.w;
    uint32_t* s_off = L.off;
    for (int i = lane; i < kWindowEntries; i += 64) s_lut[i] = lut[i];

    // Per-frame tables (strides are 0 when every frame shares one geometry).
    const ScanlineDT* scans = a.scans + static_cast<size_t>(t.frame) * a.scan_frame_stride;
    const float* depths = a.depths + static_cast<size_t>(t.frame) * a.depth_frame_stride;
    const float* ex = a.ex + static_cast<size_t>(t.frame) * a.el_frame_stride;
    const float c = a.frame_c ? load_uniform_float(a.frame_c, t.frame) : a.c;

    const ScanView sc = scan_view(scans, t.s);
    Pixel2D p;
    p.d = t.valid ? depths[t.r] : 0.0f;
    p.aDT = (((p.d * a.dt) * c) / (2.0f * a.F)) / c / a.dt;
    p.inv = 1.0f / fminf(p.aDT, sc.max_ed_x());
    p.px = sc.x() + sc.dir_x() * p.d;
    const float py = sc.dir_y() * p.d;
    p.py2 = py * py;
    const bool vertical = (sc.dir_x() == 0.0f);
    // aDT at the tile's first depth (the smallest one: the host enables the
    // far-field path only for non-negative, non-decreasing depth tables)
    const float d0 = load_uniform_float(depths, t.r0);
    const float aDT0 = (((d0 * a.dt) * c) / (2.0f * a.F)) / c / a.dt;
    const float inv_far = 1.0f / sc.max_ed_x();
    const size_t norm_slot_stride = static_cast<size_t>(a.n_rx) * a.n_depths;
    float* norm_px = a.norm_out ? a.norm_out + (static_cast<size_t>(t.frame) * (INTERP_TX ? 4 : 1) * a.n_rx + t.s) *
                                                a.n_depths + t.r
                             : nullptr;

    const RFT* rf_frame = static_cast<const RFT*>(a.rf) + static_cast<uint64_t>(t.frame) * a.rf_frame_stride;
    float s_interp = 0.0f;

    for (int k = 0; k < 4; k++) {
        const bool use = INTERP_TX ? ((sc.active_mask() >> k) & 1) != 0 : (k == sc.sel());
        if (!use) {   // uniform over the workgroup
            if (INTERP_TX && norm_px && t.valid) norm_px[k * norm_slot_stride] = 0.0f;
            continue;
        }
        const TxDT tx = sc.tx(k);
        float s_local;
        if (TESTSIG) {
            s_local = test_signal_2d(sc.x(), sc.dir_x(), sc.dir_y(), p.d, a.dt, c);
        } else {
            const RFT* rf_tx = rf_frame + static_cast<uint64_t>(tx.tx_idx) * a.n_ch * a.n_s;
            const __amdgpu_buffer_rsrc_t rsrc = rf_resource(rf_tx, a.n_ch * a.n_s * sizeof(RFT));
            const bool far = a.far_ok && aDT0 >= tx.far_thr;
            float sample = 0.0f, wsum = 0.0f;
            int nadd = 0;
            const int first = tx.first_x, last = tx.last_x;
            for (int base = first; base < last; base += kChunk2D) {
                const int cnt = min(kChunk2D, last - base);
                wave_sync_2d();
                for (int j = lane; j < cnt; j += 64) {
                    const int e = base + j;
                    uint32_t ch;
                    if (MAP) {
                        const int32_t m = a.e2c[e];
                        ch = (m < 0) ? kNotConnected : static_cast<uint32_t>(m);
                    } else {
                        ch = static_cast<uint32_t>(e) % a.n_ch;
                    }
                    const float xe = ex[e];
                    const float dx = xe - sc.x();
                    s_xe[j] = xe;
                    s_dx[j] = dx;
                    s_dx2[j] = dx * dx;
                    s_w[j] = s_lut[window_index(dx * inv_far)];
                    s_off[j] = (ch == kNotConnected) ? kNotConnected : ch * a.n_s;
                }
                wave_sync_2d();
                if (far) {
                    if (vertical)
                        elements2d<RFT, true, true, MAP>(rsrc, a.n_s, tx.t0, a.offset, p, s_xe, s_dx, s_dx2, s_w,
                                                         s_off, s_lut, cnt, sample, wsum, nadd);
                    else
                        elements2d<RFT, true, false, MAP>(rsrc, a.n_s, tx.t0, a.offset, p, s_xe, s_dx, s_dx2, s_w,
                                                          s_off, s_lut, cnt, sample, wsum, nadd);
                } else {
                    if (vertical)
                        elements2d<RFT, false, true, MAP>(rsrc, a.n_s, tx.t0, a.offset, p, s_xe, s_dx, s_dx2, s_w,
                                                          s_off, s_lut, cnt, sample, wsum, nadd);
                    else
                        elements2d<RFT, false, false, MAP>(rsrc, a.n_s, tx.t0, a.offset, p, s_xe, s_dx, s_dx2, s_w,
                                                           s_off, s_lut, cnt, sample, wsum, nadd);
                }
            }
            s_local = (nadd > 0) ? (sample / wsum) * static_cast<float>(nadd) : 0.0f;
            if (norm_px && t.valid) {
                const float g = (nadd > 0) ? static_cast<float>(nadd) / wsum : 0.0f;
                norm_px[(INTERP_TX ? k : 0) * norm_slot_stride] = INTERP_TX ? sc.txw(k) * g : g;
            }
        }
        if (INTERP_TX) s_interp += sc.txw(k) * s_local;
        else s_interp += s_local;
    }
    if (t.valid) {
        OutT* out = static_cast<OutT*>(a.out) + static_cast<uint64_t>(t.frame) * a.out_frame_stride;
        out[static_cast<uint64_t>(t.r) * a.n_rx + t.s] = clamp_cast<OutT>(s_interp);
        // A saturated pixel (NaN or beyond FLT_MAX) is constant under small
        // changes of its inputs: zero derivative.
        if (norm_px && !(fabsf(s_interp) <= FLT_MAX))
            for (int k = 0; k < (INTERP_TX ? 4 : 1); k++) norm_px[k * norm_slot_stride] = 0.0f;
    }
}

template <typename RFT, typename OutT, bool TESTSIG, bool INTERP_TX, bool MAP>
__global__ __launch_bounds__(kThreads) void das2d_kernel(const DasArgs a) {
    __shared__ Lds2D lds[kWavesPerGroup];
    const uint32_t wave = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / 64)));
    das2d_tile<RFT, OutT, TESTSIG, INTERP_TX, MAP>(a, a.lut, decode_wave_2d(a), lds[wave]);
}

// ---- 2D, several frames per wave ---------------------------------------------
//
// Frames that share one geometry and one set of parameters take the same
// decisions for a pixel: aperture membership, window weight, W and n, the
// delay, its floor and fraction and the range test depend only on the
// geometry.  One wave therefore takes one 64-depth unit of one scanline for
// FB frames at once: it evaluates an element's delay (the correctly rounded
// square root) and decisions once, then gathers and interpolates the
// element's two samples in each of the FB frames.  Every frame's arithmetic
// is the single-frame kernel's, in the same order, so the images are the
// same bits.

// Elements [0, cnt) of the current LDS chunk for one 2D pixel in FB frames.
template <typename RFT, int FB, bool FAR, bool VERT, bool MAP>
__device__ inline void elements2d_frames(const __amdgpu_buffer_rsrc_t (&rs)[FB], uint32_t n_s, float t0,
                                         float offset, const Pixel2D& p, const Lds2D& L, int cnt,
                                         float (&sample)[FB], float& wsum, int& nadd) {
    const float last = static_cast<float>(n_s - 1u);
#pragma unroll 2
    for (int j = 0; j < cnt; j++) {
        const uint32_t off = L.off[j];
        if (MAP && off == kNotConnected) continue;
        float w;
        if (FAR) {
            w = L.w[j];
        } else {
            const float dx = L.dx[j];
            if (!(fabsf(dx) <= p.aDT)) continue;
            w = L.lut[window_index(dx * p.inv)];
        }
        wsum += w;
        nadd++;
        float q;
        if (VERT) {
            q = L.dx2[j] + p.py2;
        } else {
            const float e = L.xe[j] - p.px;
            q = e * e + p.py2;
        }
        float tau = t0 + (sqrt_rn(q) + p.d);
        tau = tau + offset;
        const uint32_t ui = static_cast<uint32_t>(static_cast<int32_t>(floorf(tau)));
        const float f = __builtin_amdgcn_fractf(tau);
        const float omf = 1.0f - f;
        const uint32_t idx = off + ui;
        if constexpr (std::is_same_v<RFT, int16_t>) {
            // 0 <= tau <= N-1 (see gather()); outside it the element adds
            // w_eff * val = +-0 to the sum, which leaves it unchanged (int16
            // samples are finite).
            const bool in = (tau >= 0.0f) & (tau <= last);
            const float w_eff = in ? w : 0.0f;
            float v0[FB], v1[FB];
#pragma unroll
            for (int fb = 0; fb < FB; fb++) {     // every frame's gathers in flight at once
                v0[fb] = load_rf<RFT>(rs[fb], idx);
                v1[fb] = load_rf<RFT>(rs[fb], idx + 1u);
            }
#pragma unroll
            for (int fb = 0; fb < FB; fb++) sample[fb] += w_eff * (omf * v0[fb] + f * v1[fb]);
        } else {
            const bool inr = ui < n_s - 1u;
            const bool edge = (ui == n_s - 1u) & (f == 0.0f);
            float v0[FB], v1[FB];
#pragma unroll
            for (int fb = 0; fb < FB; fb++) {
                v0[fb] = load_rf<RFT>(rs[fb], idx);
                v1[fb] = load_rf<RFT>(rs[fb], idx + 1u);
            }
#pragma unroll
            for (int fb = 0; fb < FB; fb++) {
                const float val = edge ? v0[fb] : (omf * v0[fb] + f * v1[fb]);
                const float contrib = w * val;
                sample[fb] = (inr | edge) ? sample[fb] + contrib : sample[fb];
            }
        }
    }
}

// One wave: 64 consecutive depths of one scanline in frames f0 .. f0+FB-1
// (the last group may hold fewer frames; its missing frames re-read the last
// frame and are not written).  DelayAndSum without transmit interpolation,
// one geometry for all frames.
template <typename RFT, typename OutT, int FB, bool MAP>
__global__ __launch_bounds__(kThreads) void das2d_frames_kernel(const DasArgs a) {
    __shared__ Lds2D lds[kWavesPerGroup];
    const uint32_t wave = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / 64)));
    const int lane = static_cast<int>(threadIdx.x % 64);
    const uint32_t chunks = (a.n_depths + kWaveDepths - 1) / kWaveDepths;
    const uint32_t groups = (a.n_frames + FB - 1) / FB;
    // Units in tiles of kTileLines scanlines x kTileChunks chunks (scanline
    // fastest, then chunk), tiles ordered (scanline block, frame group, chunk
    // pair from the deepest), dealt to the XCDs as whole tiles.  Units past a
    // ragged edge (scanline >= n_rx, chunk >= chunks) are empty.
    const uint64_t unit = static_cast<uint64_t>(logical_wg_2d(kXcdTileWgs * kTileChunks)) * kWavesPerGroup + wave;
    const uint32_t sblocks = (a.n_rx + kTileLines - 1) / kTileLines;
    const uint64_t tile = unit / (kTileLines * kTileChunks);
    const uint32_t in_tile = static_cast<uint32_t>(unit % (kTileLines * kTileChunks));
    const uint32_t s = static_cast<uint32_t>(tile % sblocks) * kTileLines + in_tile % kTileLines;
    const uint64_t rest = tile / sblocks;
    const uint64_t grp = rest % groups;
    const uint64_t cr = (rest / groups) * kTileChunks + in_tile / kTileLines;   // chunk, counted from the deepest
    if (cr >= chunks || s >= a.n_rx || grp >= groups) return;   // (wave-uniform) an empty unit
    const uint32_t r0 = (chunks - 1u - static_cast<uint32_t>(cr)) * kWaveDepths;
    const uint32_t r = r0 + static_cast<uint32_t>(lane);
    const bool valid = r < a.n_depths;
    const uint32_t f0 = static_cast<uint32_t>(grp) * FB;
    const uint32_t nfr = min(static_cast<uint32_t>(FB), a.n_frames - f0);
    Lds2D& L = lds[wave];
    for (int i = lane; i < kWindowEntries; i += 64) L.lut[i] = a.lut[i];

    const ScanView sc = scan_view(a.scans, s);
    Pixel2D p;
    p.d = valid ? a.depths[r] : 0.0f;
    p.aDT = (((p.d * a.dt) * a.c) / (2.0f * a.F)) / a.c / a.dt;
    p.inv = 1.0f / fminf(p.aDT, sc.max_ed_x());
    p.px = sc.x() + sc.dir_x() * p.d;
    const float py = sc.dir_y() * p.d;
    p.py2 = py * py;
    const bool vertical = (sc.dir_x() == 0.0f);
    const float d0 = load_uniform_float(a.depths, r0);
    const float aDT0 = (((d0 * a.dt) * a.c) / (2.0f * a.F)) / a.c / a.dt;
    const float inv_far = 1.0f / sc.max_ed_x();

    float sample[FB];
#pragma unroll
    for (int fb = 0; fb < FB; fb++) sample[fb] = 0.0f;
    float wsum = 0.0f;
    int nadd = 0;
    const int k = sc.sel();
    if (k >= 0) {                                    // uniform
        const TxDT tx = sc.tx(k);
        const uint64_t tx_off = static_cast<uint64_t>(tx.tx_idx) * a.n_ch * a.n_s;
        __amdgpu_buffer_rsrc_t rs[FB];
#pragma unroll
        for (int fb = 0; fb < FB; fb++) {
            const uint32_t frame = f0 + min(static_cast<uint32_t>(fb), nfr - 1u);
            const RFT* base = static_cast<const RFT*>(a.rf) + static_cast<uint64_t>(frame) * a.rf_frame_stride + tx_off;
            rs[fb] = rf_resource(base, a.n_ch * a.n_s * sizeof(RFT));
        }
        const bool far = a.far_ok && aDT0 >= tx.far_thr;
        const int first = tx.first_x, last = tx.last_x;
        for (int base = first; base < last; base += kChunk2D) {
            const int cnt = min(kChunk2D, last - base);
            wave_sync_2d();
            for (int j = lane; j < cnt; j += 64) {
                const int e = base + j;
                uint32_t ch;
                if (MAP) {
                    const int32_t m = a.e2c[e];
                    ch = (m < 0) ? kNotConnected : static_cast<uint32_t>(m);
                } else {
                    ch = static_cast<uint32_t>(e) % a.n_ch;
                }
                const float xe = a.ex[e];
                const float dx = xe - sc.x();
                L.xe[j] = xe;
                L.dx[j] = dx;
                L.dx2[j] = dx * dx;
                L.w[j] = L.lut[window_index(dx * inv_far)];
                L.off[j] = (ch == kNotConnected) ? kNotConnected : ch * a.n_s;
            }
            wave_sync_2d();
            if (far) {
                if (vertical) elements2d_frames<RFT, FB, true, true, MAP>(rs, a.n_s, tx.t0, a.offset, p, L, cnt, sample, wsum, nadd);
                else elements2d_frames<RFT, FB, true, false, MAP>(rs, a.n_s, tx.t0, a.offset, p, L, cnt, sample, wsum, nadd);
            } else {
                if (vertical) elements2d_frames<RFT, FB, false, true, MAP>(rs, a.n_s, tx.t0, a.offset, p, L, cnt, sample, wsum, nadd);
                else elements2d_frames<RFT, FB, false, false, MAP>(rs, a.n_s, tx.t0, a.offset, p, L, cnt, sample, wsum, nadd);
            }
        }
    }
    if (!valid) return;
#pragma unroll
    for (int fb = 0; fb < FB; fb++) {
        if (static_cast<uint32_t>(fb) >= nfr) break;   // uniform
        const float v = (nadd > 0) ? (sample[fb] / wsum) * static_cast<float>(nadd) : 0.0f;
        OutT* out = static_cast<OutT*>(a.out) + static_cast<uint64_t>(f0 + fb) * a.out_frame_stride;
        out[static_cast<uint64_t>(r) * a.n_rx + s] = clamp_cast<OutT>(v);
    }
}

// ---- 3D ---------------------------------------------------------------------

struct Pixel3D {
    float d, aDT, invx, invz, px, pz, py2;
};

// One 3D element's gathers, issued before any of a batch is consumed
// (gather<RFT, true>'s loads and decisions, RxSampleBeamformerDelayAndSum.h:
// 82-106).  A skipped element (outside the aperture, unconnected channel)
// has member == false and weight 0.
template <typename RFT>
struct Gather3D {
    float v0, v1, f, w;
    bool add;    // int16: member and 0 <= tau <= N-1 (saturated floor); float: member and (inr | edge)
    bool edge;   // float: the last sample with zero fraction
};

template <typename RFT>
__device__ inline Gather3D<RFT> gather3d_issue(__amdgpu_buffer_rsrc_t rf, uint32_t off, uint32_t n_s, float tau,
                                               float w, bool member) {
    Gather3D<RFT> g;
    const float fl = fmaxf(floorf(tau), 0.0f);
    const uint32_t ui = static_cast<uint32_t>(fl);
    g.f = tau - fl;
    g.w = w;
    g.v0 = load_rf<RFT>(rf, off + ui);
    g.v1 = load_rf<RFT>(rf, off + ui + 1u);
    if constexpr (std::is_same_v<RFT, int16_t>) {
        const float last = static_cast<float>(n_s - 1u);
        g.add = member & !(tau > last) & !(tau < (n_s > 1u ? -INFINITY : 0.0f));
        g.edge = false;
    } else {
        const bool inr = ui < n_s - 1u;
        g.edge = (ui == n_s - 1u) & (g.f == 0.0f);
        g.add = member & (inr | g.edge);
    }
    return g;
}

// The element's contribution, in the single-element path's operations: for
// int16 a skipped element adds +0 (the sum is never -0: an interpolated
// int16 value is never -0, so a contribution of -0 needs w = 0, and +0 + -0
// is +0), for float RF the sum is selected unchanged.
template <typename RFT>
__device__ inline void gather3d_consume(const Gather3D<RFT>& g, float& sample) {
    if constexpr (std::is_same_v<RFT, int16_t>) {
        const float contrib = g.w * ((1.0f - g.f) * g.v0 + g.f * g.v1);
        sample += g.add ? contrib : 0.0f;
    } else {
        const float val = g.edge ? g.v0 : ((1.0f - g.f) * g.v0 + g.f * g.v1);
        const float contrib = g.w * val;
        sample = g.add ? sample + contrib : sample;
    }
}

// Elements [0, cnt) of the current LDS chunk for one 3D pixel
// (RxSampleBeamformerDelayAndSum.h:55-111, x outer, y inner).  Elements go in
// batches of kBatch3D: every element of a batch computes its delay and
// issues its two gathers before the first is consumed, so a wave keeps
// 2 x kBatch3D loads in flight instead of waiting out each element's memory
// latency.  Weights, counts and samples are still added in element order.
constexpr int kBatch3D = 4;   // 2: 5.63 ms, 4: 4.9-5.0 ms, 8: 5.09 ms, 16: 5.27 ms per C1 volume

// Per-row values of the element loop: elements come x outer, y inner, so
// consecutive elements of one x column share xe, hence (xe - px)^2 and the x
// window index.  They are recomputed whenever the staged xe changes (a
// wave-uniform test), so the values are the same operations on the same
// inputs as per element, for any element layout.
struct Row3D {
    uint32_t xe_bits;
    float ax2;
    int ix64;    // x window index * 64 (near field)
};

template <typename RFT, bool FAR, bool MAP>
__device__ inline Gather3D<RFT> element3d_issue(__amdgpu_buffer_rsrc_t rf, uint32_t n_s, float t0, float offset,
                                                const Pixel3D& p, const float* s_xe, const float* s_ze,
                                                const float* s_dx, const float* s_dz, const float* s_r2,
                                                const float* s_w, const uint32_t* s_off, const float* s_w2, int j,
                                                Row3D& row, float& wsum, int& nadd) {
    const float xe = s_xe[j];
    const uint32_t xe_bits = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(__float_as_int(xe)));
    if (xe_bits != row.xe_bits) {                    // uniform: a new x column
        row.xe_bits = xe_bits;
        const float ax = xe - p.px;
        row.ax2 = ax * ax;
        if (!FAR) row.ix64 = window_index(s_dx[j] * p.invx) * kWindowEntries;
    }
    const uint32_t off = s_off[j];
    bool member = !MAP || off != kNotConnected;
    float w;
    if (FAR) {
        w = member ? s_w[j] : 0.0f;
    } else {
        member = member && (s_r2[j] <= p.aDT);
        if (member) {
            const int iz = window_index(s_dz[j] * p.invz);
            w = s_w2[row.ix64 + iz];
        } else {
            w = 0.0f;
        }
    }
    // a skipped element adds +0 to wsum (never -0: window weights are >= 0) and 0 to n
    wsum += w;
    nadd += member ? 1 : 0;
    const float az = s_ze[j] - p.pz;
    float tau = t0 + (sqrt_rn((row.ax2 + az * az) + p.py2) + p.d);
    tau = tau + offset;
    return gather3d_issue<RFT>(rf, off, n_s, tau, w, member);
}

template <typename RFT, bool FAR, bool MAP>
__device__ inline void elements3d(__amdgpu_buffer_rsrc_t rf, uint32_t n_s, float t0, float offset, const Pixel3D& p,
                                  const float* s_xe, const float* s_ze, const float* s_dx, const float* s_dz,
                                  const float* s_r2, const float* s_w, const uint32_t* s_off, const float* s_w2,
                                  int cnt, float& sample, float& wsum, int& nadd) {
    // xe is never the NaN pattern 0xffffffff, so the first element starts a row
    Row3D row{0xffffffffu, 0.0f, 0};
    int j = 0;
    for (; j + kBatch3D <= cnt; j += kBatch3D) {
        if (!FAR) {
            // near field: a batch that no lane of the wave has in its
            // aperture adds nothing (+0 weights, no samples) and is skipped
            bool any = false;
#pragma unroll
            for (int u = 0; u < kBatch3D; u++)
                any |= (s_r2[j + u] <= p.aDT) && (!MAP || s_off[j + u] != kNotConnected);
            if (__builtin_amdgcn_ballot_w64(any) == 0) continue;   // uniform
        }
        Gather3D<RFT> g[kBatch3D];
#pragma unroll
        for (int u = 0; u < kBatch3D; u++)
            g[u] = element3d_issue<RFT, FAR, MAP>(rf, n_s, t0, offset, p, s_xe, s_ze, s_dx, s_dz, s_r2, s_w, s_off,
                                                  s_w2, j + u, row, wsum, nadd);
#pragma unroll
        for (int u = 0; u < kBatch3D; u++) gather3d_consume<RFT>(g[u], sample);
    }
    for (; j < cnt; j++) {
        const Gather3D<RFT> g = element3d_issue<RFT, FAR, MAP>(rf, n_s, t0, offset, p, s_xe, s_ze, s_dx, s_dz, s_r2,
                                                               s_w, s_off, s_w2, j, row, wsum, nadd);
        gather3d_consume<RFT>(g, sample);
    }
}

// RxSampleBeamformerTestSignal::sampleBeamform3D (RxSampleBeamformerTestSignal.h:25-85).
__device__ inline float test_signal_3d(const ScanView& sc, float d, float dt, float c) {
    const float px = ((sc.x() + sc.dir_x() * d) * dt) * c;
    const float py = ((sc.dir_y() * d) * dt) * c;
    const float pz = ((sc.z() + sc.dir_z() * d) * dt) * c;
    float sample = 0.0f;
    for (int k = -3; k <= 3; k++) {
        const float ex = px - static_cast<float>(k) * 6.0f, ey = py - 30.0f, ez = 0.0f;
        if (sqrtf(ex * ex + ey * ey + ez * ez) <= 2.0f) sample = 1000.0f;
    }
    for (int k = -3; k <= 3; k++) {
        const float ex = 0.0f, ey = py - 30.0f, ez = pz - static_cast<float>(k) * 6.0f;
        if (sqrtf(ex * ex + ey * ey + ez * ez) <= 2.0f) sample = 1000.0f;
    }
    return sample;
}

template <typename RFT, typename OutT, bool TESTSIG, bool INTERP_TX, bool MAP>
__global__ __launch_bounds__(kThreads) void das3d_kernel(const DasArgs a) {
    __shared__ float s_w2[kW2Entries];   // sqrt(LUT[ix] * LUT[iz]) at ix * 64 + iz
    __shared__ float s_xe[kChunk3D];
    __shared__ float s_ze[kChunk3D];
    __shared__ float s_dx[kChunk3D];
    __shared__ float s_dz[kChunk3D];
    __shared__ float s_r2[kChunk3D];
    __shared__ float s_w[kChunk3D];
    __shared__ uint32_t s_off[kChunk3D];

    // All depth tiles of a scanline form one XCD tile (logical_wg_2d): the
    // RF lines two depth tiles share at their boundary are fetched into one
    // L2, and the XCDs take whole scanlines.  The grid is padded to whole
    // rounds; a padded workgroup has no tile.
    const uint32_t lb = logical_wg_2d(a.depth_tiles);
    if (lb >= a.depth_tiles * a.n_rx * a.n_frames) return;   // whole workgroup
    const Tile t = decode_tile(a, lb);
    // computeWindow3DShared (RxBeamformerCommon.h:54-61) as a 64x64 table:
    // the product and the square root are the same IEEE operations.
    for (int i = threadIdx.x; i < kWindowEntries * kWindowEntries; i += kThreads) {
        const int ix = i / kWindowEntries, iz = i % kWindowEntries;
        s_w2[i] = sqrtf(a.lut[ix] * a.lut[iz]);
    }

    const ScanView sc = scan_view(a.scans, t.s);
    Pixel3D p;
    p.d = t.valid ? a.depths[t.r] : 0.0f;
    {
        const float ap = (((p.d * a.dt) * a.c) / (2.0f * a.F)) / a.c / a.dt;
        p.aDT = ap * ap;
        const float sq = sqrtf(p.aDT);
        p.invx = 1.0f / fminf(sq, sc.max_ed_x());
        p.invz = 1.0f / fminf(sq, sc.max_ed_y());
    }
    p.px = sc.x() + sc.dir_x() * p.d;
    p.pz = sc.z() + sc.dir_z() * p.d;
    {
        const float py = sc.dir_y() * p.d;
        p.py2 = py * py;
    }
    // Far-field test per wave, at the wave's first (smallest) depth: a wave
    // whose 64 depths all see the saturated aperture takes the far path even
    // when the workgroup's first wave does not.
    const uint32_t wave = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / 64)));
    const float d0 = load_uniform_float(a.depths, min(t.r0 + wave * 64u, a.n_depths - 1u));
    const float ap0 = (((d0 * a.dt) * a.c) / (2.0f * a.F)) / a.c / a.dt;
    const float aDT0 = ap0 * ap0;
    const float sq0 = sqrtf(aDT0);
    const float invx_far = 1.0f / sc.max_ed_x(), invz_far = 1.0f / sc.max_ed_y();

    const RFT* rf_frame = static_cast<const RFT*>(a.rf) + static_cast<uint64_t>(t.frame) * a.rf_frame_stride;
    float s_interp = 0.0f;

    for (int k = 0; k < 4; k++) {
        const bool use = INTERP_TX ? ((sc.active_mask() >> k) & 1) != 0 : (k == sc.sel());
        if (!use) continue;
        const TxDT tx = sc.tx(k);
        float s_local;
        if (TESTSIG) {
            s_local = test_signal_3d(sc, p.d, a.dt, a.c);
        } else {
            const RFT* rf_tx = rf_frame + static_cast<uint64_t>(tx.tx_idx) * a.n_ch * a.n_s;
            const __amdgpu_buffer_rsrc_t rsrc = rf_resource(rf_tx, a.n_ch * a.n_s * sizeof(RFT));
            const bool far = a.far_ok && aDT0 >= tx.far_thr && sq0 >= sc.max_ed_x() && sq0 >= sc.max_ed_y();
            float sample = 0.0f, wsum = 0.0f;
            int nadd = 0;
            const int ny = (tx.last_y > tx.first_y) ? tx.last_y - tx.first_y : 0;
            const int nx = (tx.last_x > tx.first_x) ? tx.last_x - tx.first_x : 0;
            const int total = nx * ny;
            // element j of the aperture: ex = first_x + j / ny, ey = first_y + j % ny
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
                        const int32_t m = a.e2c[e];
                        ch = (m < 0) ? kNotConnected : static_cast<uint32_t>(m);
                    } else {
                        ch = e % a.n_ch;
                    }
                    const float xe = a.ex[e], ze = a.ey[e];
                    const float dx = xe - sc.x(), dz = ze - sc.z();
                    s_xe[j] = xe;
                    s_ze[j] = ze;
                    s_dx[j] = dx;
                    s_dz[j] = dz;
                    s_r2[j] = dx * dx + dz * dz;
                    s_w[j] = s_w2[window_index(dx * invx_far) * kWindowEntries + window_index(dz * invz_far)];
                    s_off[j] = (ch == kNotConnected) ? kNotConnected : ch * a.n_s;
                }
                __syncthreads();
                if (far)
                    elements3d<RFT, true, MAP>(rsrc, a.n_s, tx.t0, a.offset, p, s_xe, s_ze, s_dx, s_dz, s_r2, s_w,
                                               s_off, s_w2, cnt, sample, wsum, nadd);
                else
                    elements3d<RFT, false, MAP>(rsrc, a.n_s, tx.t0, a.offset, p, s_xe, s_ze, s_dx, s_dz, s_r2, s_w,
                                                s_off, s_w2, cnt, sample, wsum, nadd);
            }
            s_local = (nadd > 0) ? (sample / wsum) * static_cast<float>(nadd) : 0.0f;
        }
        if (INTERP_TX) s_interp += sc.txw(k) * s_local;
        else s_interp += s_local;
    }
    if (t.valid) {
        OutT* out = static_cast<OutT*>(a.out) + static_cast<uint64_t>(t.frame) * a.out_frame_stride;
        out[static_cast<uint64_t>(t.r) * a.n_rx + t.s] = clamp_cast<OutT>(s_interp);
    }
}

// ---- RawDelay (2D) -----------------------------------------------------------

// rxDelayDTSPACEKernel + sampleDelay2D (src/SupraLib/Beamformer/RawDelay.cu:221-312,
// 404-468): for the highest-weight transmit, every aperture element's delayed,
// apodised sample times n/W goes to out[(s*N_ch + j)*N_d + r], j = e - first;
// slots j >= last - first are written as 0.  Lanes = depths, so each channel
// row of the output tile is one coalesced 256-sample store.
template <typename RFT, typename OutT>
__global__ __launch_bounds__(kThreads) void rawdelay2d_kernel(const DasArgs a) {
    __shared__ float s_lut[kWindowEntries];
    __shared__ float s_xe[kChunk2D];
    __shared__ float s_dx[kChunk2D];
    __shared__ float s_dx2[kChunk2D];
    __shared__ uint32_t s_off[kChunk2D];

    const Tile t = decode_tile(a);
    for (int i = threadIdx.x; i < kWindowEntries; i += kThreads) s_lut[i] = a.lut[i];
    const ScanView sc = scan_view(a.scans, t.s);
    Pixel2D p;
    p.d = t.valid ? a.depths[t.r] : 0.0f;
    p.aDT = (((p.d * a.dt) * a.c) / (2.0f * a.F)) / a.c / a.dt;
    p.inv = 1.0f / fminf(p.aDT, sc.max_ed_x());
    p.px = sc.x() + sc.dir_x() * p.d;
    const float py = sc.dir_y() * p.d;
    p.py2 = py * py;
    const bool vertical = (sc.dir_x() == 0.0f);
    OutT* out = static_cast<OutT*>(a.out) + static_cast<uint64_t>(t.frame) * a.out_frame_stride +
                static_cast<uint64_t>(t.s) * a.n_ch * a.n_depths;
    const int k = sc.sel();
    int written = 0;
    if (k >= 0) {
        const TxDT tx = sc.tx(k);
        const RFT* rf_tx = static_cast<const RFT*>(a.rf) + static_cast<uint64_t>(t.frame) * a.rf_frame_stride +
                           static_cast<uint64_t>(tx.tx_idx) * a.n_ch * a.n_s;
        const __amdgpu_buffer_rsrc_t rsrc = rf_resource(rf_tx, a.n_ch * a.n_s * sizeof(RFT));
        const int first = tx.first_x, last = tx.last_x;
        // pass 1: aperture weight sum and count
        float wsum = 0.0f;
        int nadd = 0;
        for (int base = first; base < last; base += kChunk2D) {
            const int cnt = min(kChunk2D, last - base);
            __syncthreads();
            for (int j = threadIdx.x; j < cnt; j += kThreads) s_dx[j] = a.ex[base + j] - sc.x();
            __syncthreads();
            for (int j = 0; j < cnt; j++) {
                const float dx = s_dx[j];
                if (fabsf(dx) <= p.aDT) {
                    wsum += s_lut[window_index(dx * p.inv)];
                    nadd++;
                }
            }
        }
        float scale = 1.0f;
        if (nadd > 0) scale = 1.0f / wsum * static_cast<float>(nadd);
        // pass 2: one output channel per aperture element
        for (int base = first; base < last; base += kChunk2D) {
            const int cnt = min(kChunk2D, last - base);
            __syncthreads();
            for (int j = threadIdx.x; j < cnt; j += kThreads) {
                const int e = base + j;
                const float xe = a.ex[e];
                const float dx = xe - sc.x();
                s_xe[j] = xe;
                s_dx[j] = dx;
                s_dx2[j] = dx * dx;
                s_off[j] = (static_cast<uint32_t>(e) % a.n_ch) * a.n_s;
            }
            __syncthreads();
            for (int j = 0; j < cnt; j++) {
                float sample = 0.0f;
                const float dx = s_dx[j];
                if (fabsf(dx) <= p.aDT) {
                    const float w = s_lut[window_index(dx * p.inv)];
                    float q;
                    if (vertical) {
                        q = s_dx2[j] + p.py2;
                    } else {
                        const float e = s_xe[j] - p.px;
                        q = e * e + p.py2;
                    }
                    const float tau = tx.t0 + (sqrt_rn(q) + p.d);
                    gather<RFT, false>(rsrc, s_off[j], a.n_s, tau, w, sample);
                }
                if (t.valid) out[static_cast<uint64_t>(base - first + j) * a.n_depths + t.r] = clamp_cast<OutT>(sample * scale);
            }
        }
        written = last > first ? last - first : 0;
    }
    if (t.valid)
        for (uint32_t j = written; j < a.n_ch; j++) out[static_cast<uint64_t>(j) * a.n_depths + t.r] = clamp_cast<OutT>(0.0f);
}

// rxDelayDTSPACE3DKernel + sampleDelay3D (RawDelay.cu:126-220, 315-402): for
// the highest-weight transmit, the delayed, apodised sample times n/W of
// aperture element (ex, ey) goes to slot (ex - first.x) + (ey - first.y) *
// layout.x of out[(s*N_ch + slot)*N_d + r].  Every slot is written once per
// pixel: slots no aperture element maps to, elements outside the F-number
// aperture and delays out of range give 0.  Lanes = depths; the slots are
// staged in LDS 128 at a time.
template <typename RFT, typename OutT>
__global__ __launch_bounds__(kThreads) void rawdelay3d_kernel(const DasArgs a) {
    __shared__ float s_lut[kWindowEntries];
    __shared__ float s_xe[kChunk3D];
    __shared__ float s_ze[kChunk3D];
    __shared__ uint32_t s_off[kChunk3D];      // channel offset; kNotConnected: no element in this slot

    const Tile t = decode_tile(a);
    for (int i = threadIdx.x; i < kWindowEntries; i += kThreads) s_lut[i] = a.lut[i];
    const ScanView sc = scan_view(a.scans, t.s);
    const float d = t.valid ? a.depths[t.r] : 0.0f;
    const float t_ap = (((d * a.dt) * a.c) / (2.0f * a.F)) / a.c / a.dt;
    const float aDT = t_ap * t_ap;
    const float sq = sqrtf(aDT);
    const float invx = 1.0f / fminf(sq, sc.max_ed_x());
    const float invz = 1.0f / fminf(sq, sc.max_ed_y());
    const float x = sc.x(), z = sc.z();
    const float px = x + sc.dir_x() * d, pz = z + sc.dir_z() * d;
    const float py = sc.dir_y() * d;
    const float py2 = py * py;
    OutT* out = static_cast<OutT*>(a.out) + static_cast<uint64_t>(t.s) * a.n_ch * a.n_depths;
    const int k = sc.sel();
    if (k < 0) {
        __syncthreads();
        if (t.valid)
            for (uint32_t j = 0; j < a.n_ch; j++) out[static_cast<uint64_t>(j) * a.n_depths + t.r] = clamp_cast<OutT>(0.0f);
        return;
    }
    const TxDT tx = sc.tx(k);
    const uint32_t fx = tx.first_x, fy = tx.first_y;
    const uint32_t wx = tx.last_x > fx ? tx.last_x - fx : 0u, wy = tx.last_y > fy ? tx.last_y - fy : 0u;
    const uint32_t lx = a.layout_x;
    const RFT* rf_tx = static_cast<const RFT*>(a.rf) + static_cast<uint64_t>(tx.tx_idx) * a.n_ch * a.n_s;
    const __amdgpu_buffer_rsrc_t rsrc = rf_resource(rf_tx, a.n_ch * a.n_s * sizeof(RFT));
    auto weight = [&](float xe, float ze, bool& member) {
        const float dxs = xe - x, dzs = ze - z;
        member = (dxs * dxs + dzs * dzs) <= aDT;
        return sqrtf(s_lut[window_index(dxs * invx)] * s_lut[window_index(dzs * invz)]);
    };
    // pass 1: n and W over the aperture, elements in x-outer / y-inner order
    float wsum = 0.0f;
    int nadd = 0;
    const uint32_t n_ap = wx * wy;
    for (uint32_t base = 0; base < n_ap; base += kChunk3D) {
        const uint32_t cnt = min(static_cast<uint32_t>(kChunk3D), n_ap - base);
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < cnt; j += kThreads) {
            const uint32_t q = base + j;
            const uint32_t e = (fx + q / wy) + (fy + q % wy) * lx;
            s_xe[j] = a.ex[e];
            s_ze[j] = a.ey[e];
        }
        __syncthreads();
        for (uint32_t j = 0; j < cnt; j++) {
            bool m;
            const float w = weight(s_xe[j], s_ze[j], m);
            if (m) {
                wsum += w;
                nadd++;
            }
        }
    }
    float scale = 1.0f;
    if (nadd > 0) scale = 1.0f / wsum * static_cast<float>(nadd);
    // pass 2: every output slot once
    for (uint32_t base = 0; base < a.n_ch; base += kChunk3D) {
        const uint32_t cnt = min(static_cast<uint32_t>(kChunk3D), a.n_ch - base);
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < cnt; j += kThreads) {
            const uint32_t slot = base + j;
            const uint32_t ox = slot % lx, oy = slot / lx;
            if (ox < wx && oy < wy) {
                const uint32_t e = (fx + ox) + (fy + oy) * lx;
                s_xe[j] = a.ex[e];
                s_ze[j] = a.ey[e];
                s_off[j] = (e % a.n_ch) * a.n_s;
            } else {
                s_off[j] = kNotConnected;
            }
        }
        __syncthreads();
        for (uint32_t j = 0; j < cnt; j++) {
            float sample = 0.0f;
            const uint32_t off = s_off[j];
            if (off != kNotConnected) {
                bool m;
                const float w = weight(s_xe[j], s_ze[j], m);
                if (m) {
                    const float ax = s_xe[j] - px, az = s_ze[j] - pz;
                    const float tau = tx.t0 + (sqrt_rn((ax * ax + az * az) + py2) + d);
                    gather<RFT, true>(rsrc, off, a.n_s, tau, w, sample);
                }
            }
            if (t.valid) out[static_cast<uint64_t>(base + j) * a.n_depths + t.r] = clamp_cast<OutT>(sample * scale);
        }
    }
}

// ---- persistent work queue over many 2D frames -----------------------------

template <typename RFT, typename OutT>
__device__ inline void das2d_batch_tile(const DasBatchFrame& fr, const DasArgs& a, const Tile& t, Lds2D& L) {
    if (fr.interp) {
        if (fr.map) das2d_tile<RFT, OutT, false, true, true>(a, fr.a.lut, t, L);
        else das2d_tile<RFT, OutT, false, true, false>(a, fr.a.lut, t, L);
    } else {
        if (fr.map) das2d_tile<RFT, OutT, false, false, true>(a, fr.a.lut, t, L);
        else das2d_tile<RFT, OutT, false, false, false>(a, fr.a.lut, t, L);
    }
}

// Each workgroup takes the next (frame, scanline, depth tile) item until the
// queue is empty; the last item of a frame records the frame's completion
// time.  Every workgroup leaves the loop once the queue index passes n_items.
__global__ __launch_bounds__(kThreads) void das2d_batch_kernel(const DasBatchArgs b) {
    __shared__ Lds2D lds_waves[kWavesPerGroup];
    __shared__ uint32_t s_item;
    const uint32_t wave = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / 64)));
    Lds2D& lds = lds_waves[wave];
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = atomicAdd(b.sync.head, 1u);
        __syncthreads();
        const uint32_t item = s_item;
        if (item >= b.n_items) break;
        const uint32_t code = b.items[item];
        const DasBatchFrame& fr = b.frames[code >> 24];
        const DasArgs a = fr.a;     // scalars of the frame (the LUT is read from fr)
        const uint32_t local = code & 0xffffffu;
        Tile t;
        const uint32_t tile = local % a.depth_tiles;
        t.s = local / a.depth_tiles;
        t.frame = 0;
        t.r0 = tile * kThreads + wave * kWaveDepths;   // each wave: 64 depths of the item's 256
        t.r = t.r0 + threadIdx.x % 64;
        t.valid = t.r < a.n_depths;
        if (fr.rf_dtype == SUPRA_INT16) {
            if (fr.out_dtype == SUPRA_FLOAT32) das2d_batch_tile<int16_t, float>(fr, a, t, lds);
            else das2d_batch_tile<int16_t, int16_t>(fr, a, t, lds);
        } else {
            if (fr.out_dtype == SUPRA_FLOAT32) das2d_batch_tile<float, float>(fr, a, t, lds);
            else das2d_batch_tile<float, int16_t>(fr, a, t, lds);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence();
            if (atomicAdd(&b.sync.done[fr.frame], 1u) + 1u == fr.n_items) b.sync.stamp[fr.frame] = wall_clock64();
        }
    }
}

}  // namespace

// 2D grids are padded to whole rounds of XCD tiles (logical_wg_2d).
static uint64_t pad_2d_grid(uint64_t blocks) {
    const uint64_t round = static_cast<uint64_t>(kXcds) * kXcdTileWgs;
    return (blocks + round - 1) / round * round;
}

template <typename RFT, typename OutT>
static hipError_t launch_typed(const DasArgs& a, bool is3d, bool testsig, bool interp_tx, bool map,
                               hipStream_t stream) {
    const uint64_t units2d = static_cast<uint64_t>((a.n_depths + kWaveDepths - 1) / kWaveDepths) * a.n_rx * a.n_frames;
    const uint64_t round3d = static_cast<uint64_t>(kXcds) * a.depth_tiles;   // whole rounds of scanline tiles
    const uint64_t blocks = is3d ? (static_cast<uint64_t>(a.depth_tiles) * a.n_rx * a.n_frames + round3d - 1) / round3d * round3d
                                 : pad_2d_grid((units2d + kWavesPerGroup - 1) / kWavesPerGroup);
    dim3 grid(static_cast<uint32_t>(blocks)), block(kThreads);
#define SUPRA_DAS_DISPATCH(KERNEL)                                                                                 \
    do {                                                                                                            \
        if (testsig) {                                                                                              \
            if (interp_tx) hipLaunchKernelGGL((KERNEL<RFT, OutT, true, true, false>), grid, block, 0, stream, a);   \
            else hipLaunchKernelGGL((KERNEL<RFT, OutT, true, false, false>), grid, block, 0, stream, a);            \
        } else if (map) {                                                                                           \
            if (interp_tx) hipLaunchKernelGGL((KERNEL<RFT, OutT, false, true, true>), grid, block, 0, stream, a);   \
            else hipLaunchKernelGGL((KERNEL<RFT, OutT, false, false, true>), grid, block, 0, stream, a);            \
        } else {                                                                                                    \
            if (interp_tx) hipLaunchKernelGGL((KERNEL<RFT, OutT, false, true, false>), grid, block, 0, stream, a);  \
            else hipLaunchKernelGGL((KERNEL<RFT, OutT, false, false, false>), grid, block, 0, stream, a);           \
        }                                                                                                           \
    } while (0)
    if (is3d) {
        SUPRA_DAS_DISPATCH(das3d_kernel);
    } else if (!testsig && !interp_tx && a.n_frames >= 2 && a.scan_frame_stride == 0 && a.depth_frame_stride == 0 &&
               a.el_frame_stride == 0 && a.frame_c == nullptr && a.norm_out == nullptr) {
        // frames sharing one geometry: kFramesPerWave frames per wave
        constexpr int FB = kFramesPerWave;
        const uint64_t chunks = (a.n_depths + kWaveDepths - 1) / kWaveDepths;
        const uint64_t units = static_cast<uint64_t>((a.n_rx + kTileLines - 1) / kTileLines) * kTileLines *
                               ((a.n_frames + FB - 1) / FB) * ((chunks + kTileChunks - 1) / kTileChunks) * kTileChunks;
        const uint64_t round = static_cast<uint64_t>(kXcds) * kXcdTileWgs * kTileChunks;
        dim3 g2(static_cast<uint32_t>(((units + kWavesPerGroup - 1) / kWavesPerGroup + round - 1) / round * round));
        if (map) hipLaunchKernelGGL((das2d_frames_kernel<RFT, OutT, FB, true>), g2, block, 0, stream, a);
        else hipLaunchKernelGGL((das2d_frames_kernel<RFT, OutT, FB, false>), g2, block, 0, stream, a);
    } else {
        SUPRA_DAS_DISPATCH(das2d_kernel);
    }
#undef SUPRA_DAS_DISPATCH
    return hipGetLastError();
}

uint32_t das_depth_tile() { return kThreads; }

size_t das_batch_frame_bytes() { return sizeof(DasBatchFrame); }

hipError_t launch_das_batch(const DasBatchArgs& b, uint32_t workgroups, hipStream_t stream) {
    hipLaunchKernelGGL(das2d_batch_kernel, dim3(workgroups), dim3(kThreads), 0, stream, b);
    return hipGetLastError();
}

hipError_t launch_rawdelay(const DasArgs& a, int rf_dtype, int out_dtype, bool is3d, hipStream_t stream) {
    const uint64_t blocks = static_cast<uint64_t>(a.depth_tiles) * a.n_rx * a.n_frames;
    dim3 grid(static_cast<uint32_t>(blocks)), block(kThreads);
    if (is3d) {
        if (rf_dtype == SUPRA_INT16) {
            if (out_dtype == SUPRA_FLOAT32) hipLaunchKernelGGL((rawdelay3d_kernel<int16_t, float>), grid, block, 0, stream, a);
            else hipLaunchKernelGGL((rawdelay3d_kernel<int16_t, int16_t>), grid, block, 0, stream, a);
        } else {
            if (out_dtype == SUPRA_FLOAT32) hipLaunchKernelGGL((rawdelay3d_kernel<float, float>), grid, block, 0, stream, a);
            else hipLaunchKernelGGL((rawdelay3d_kernel<float, int16_t>), grid, block, 0, stream, a);
        }
        return hipGetLastError();
    }
    if (rf_dtype == SUPRA_INT16) {
        if (out_dtype == SUPRA_FLOAT32) hipLaunchKernelGGL((rawdelay2d_kernel<int16_t, float>), grid, block, 0, stream, a);
        else hipLaunchKernelGGL((rawdelay2d_kernel<int16_t, int16_t>), grid, block, 0, stream, a);
    } else {
        if (out_dtype == SUPRA_FLOAT32) hipLaunchKernelGGL((rawdelay2d_kernel<float, float>), grid, block, 0, stream, a);
        else hipLaunchKernelGGL((rawdelay2d_kernel<float, int16_t>), grid, block, 0, stream, a);
    }
    return hipGetLastError();
}

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
