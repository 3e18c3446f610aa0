// Engine context and the DAS entry points of the C ABI.
//
// Stands in for RxBeamformerCuda (src/SupraLib/Beamformer/RxBeamformerCuda.cu:24-107,
// 539-663): the constructor keeps the receive geometry, and every call whose
// sampling interval or speed of sound differs from the previous one rebuilds
// the dt-space tables.  Unlike the reference, which rescales its float
// buffers by old/new factors on each change (:60-102, so repeated changes
// accumulate rounding), the tables are always rebuilt from the millimetre
// geometry with the reference's first-call factors 1/(c*dt) and 1/dt.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "bf_internal.hpp"
#include "das_launch.hpp"
#include "mv_launch.hpp"
#include "status.hpp"
#include "supra_bf.h"

using namespace supra_amd;

#include "bf_ctx.hpp"


namespace supra_amd {

int hip_fail(hipError_t e, const char* what) {
    return fail(SUPRA_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

// RxBeamformerCuda::convertToDtSpace (RxBeamformerCuda.cu:58-107), first-call
// factors, from the millimetre geometry; then the float casts of the kernels.
void build_dt_tables(supra_bf_ctx* ctx, double dt, double c) {
    const double factor = 1 / 1.0 / (c * dt);
    const double factor_time = 1 / 1.0 / dt;
    ctx->h_depths.resize(ctx->n_depths);
    ctx->far_ok = true;
    for (uint32_t i = 0; i < ctx->n_depths; i++) {
        ctx->h_depths[i] = static_cast<float>(ctx->depths_mm[i] * factor);
        if (!(ctx->h_depths[i] >= 0.0f) || (i > 0 && !(ctx->h_depths[i] >= ctx->h_depths[i - 1])))
            ctx->far_ok = false;
    }
    ctx->h_ex.resize(ctx->n_el);
    ctx->h_ey.resize(ctx->n_el);
    for (uint32_t i = 0; i < ctx->n_el; i++) {
        ctx->h_ex[i] = static_cast<float>(ctx->ex_mm[i] * factor);
        ctx->h_ey[i] = static_cast<float>(ctx->ey_mm[i] * factor);
    }
    auto connected = [&](uint32_t e) { return ctx->e2c.empty() || ctx->e2c[e] >= 0; };
    ctx->h_scans.resize(ctx->n_rx);
    for (uint32_t i = 0; i < ctx->n_rx; i++) {
        const supra_scanline_rx& s = ctx->scan_mm[i];
        ScanlineDT& o = ctx->h_scans[i];
        std::memset(&o, 0, sizeof(o));
        o.x = static_cast<float>(s.position[0] * factor);
        o.z = static_cast<float>(s.position[2] * factor);
        o.dir_x = static_cast<float>(s.direction[0]);
        o.dir_y = static_cast<float>(s.direction[1]);
        o.dir_z = static_cast<float>(s.direction[2]);
        o.max_ed_x = static_cast<float>(s.max_element_distance[0] * factor);
        o.max_ed_y = static_cast<float>(s.max_element_distance[1] * factor);
        scanline_selection(s, &o.sel, &o.active_mask, o.txw);
        for (int k = 0; k < 4; k++) {
            TxDT& t = o.tx[k];
            t.first_x = s.tx[k].first_active_element[0];
            t.first_y = s.tx[k].first_active_element[1];
            t.last_x = s.tx[k].last_active_element[0];
            t.last_y = s.tx[k].last_active_element[1];
            t.tx_idx = s.tx[k].tx_scanline_idx;
            t.t0 = static_cast<float>(s.tx[k].initial_delay * factor_time);
            // Same single-precision operations as the kernels' aperture test.
            float thr = 0.0f;
            if (ctx->is3d) {
                for (uint32_t ex = t.first_x; ex < t.last_x; ex++)
                    for (uint32_t ey = t.first_y; ey < t.last_y; ey++) {
                        const uint32_t e = ex + ey * ctx->el_layout[0];
                        if (!connected(e)) continue;
                        const float dx = ctx->h_ex[e] - o.x, dz = ctx->h_ey[e] - o.z;
                        thr = std::max(thr, dx * dx + dz * dz);
                    }
            } else {
                for (uint32_t e = t.first_x; e < t.last_x; e++)
                    if (connected(e)) thr = std::max(thr, std::fabs(ctx->h_ex[e] - o.x));
                thr = std::max(thr, o.max_ed_x);
            }
            t.far_thr = thr;
        }
    }
}

int ensure_dt_tables(supra_bf_ctx* ctx, double dt, double c, hipStream_t stream) {
    if (ctx->cached_dt == dt && ctx->cached_c == c) return SUPRA_OK;
    build_dt_tables(ctx, dt, c);
    hipError_t e;
    // Kernels of earlier calls (possibly on other streams) may still read the
    // tables: the upload waits for the last of them.
    if (ctx->tables_used_recorded && (e = hipStreamWaitEvent(stream, ctx->tables_used, 0)) != hipSuccess)
        return hip_fail(e, "geometry upload: wait for earlier kernels");
    if ((e = hipMemcpyAsync(ctx->d_scans, ctx->h_scans.data(), ctx->n_rx * sizeof(ScanlineDT), hipMemcpyHostToDevice,
                            stream)) != hipSuccess)
        return hip_fail(e, "upload scanlines");
    if ((e = hipMemcpyAsync(ctx->d_depths, ctx->h_depths.data(), ctx->n_depths * sizeof(float), hipMemcpyHostToDevice,
                            stream)) != hipSuccess)
        return hip_fail(e, "upload depths");
    if ((e = hipMemcpyAsync(ctx->d_ex, ctx->h_ex.data(), ctx->n_el * sizeof(float), hipMemcpyHostToDevice, stream)) !=
        hipSuccess)
        return hip_fail(e, "upload element x");
    if ((e = hipMemcpyAsync(ctx->d_ey, ctx->h_ey.data(), ctx->n_el * sizeof(float), hipMemcpyHostToDevice, stream)) !=
        hipSuccess)
        return hip_fail(e, "upload element y");
    // The host tables are rewritten on the next change: wait for the copies.
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "geometry upload");
    ctx->cached_dt = dt;
    ctx->cached_c = c;
    return SUPRA_OK;
}

// Marks the end of a launch that reads the context's dt-space tables.
int tables_used_on(supra_bf_ctx* ctx, hipStream_t stream) {
    hipError_t e = hipEventRecord(ctx->tables_used, stream);
    if (e != hipSuccess) return hip_fail(e, "record table use");
    ctx->tables_used_recorded = true;
    return SUPRA_OK;
}

int check_device(const supra_bf_ctx* ctx) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev != ctx->device)
        return fail(SUPRA_ERR_INVALID_ARG, "context was created on device %d, current device is %d", ctx->device, dev);
    return SUPRA_OK;
}

}  // namespace

// Device scratch of supra_das_forward_c_noalloc: the per-frame dt-space
// tables are built in this file-scope array (one copy per GPU, held until
// the process ends) instead of in memory the library allocates.  Host code
// never uses the array's name as an address: supra_bf_create looks up each
// device's copy once, with that device current, and keeps the pointer.
static __device__ __align__(256) unsigned char scratch[SUPRA_NOALLOC_SCRATCH_BYTES];

namespace {

int ensure_mm_geometry(supra_bf_ctx* ctx);

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;   // per device; nullptr until looked up

int noalloc_scratch(int device, void** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0) return fail(SUPRA_ERR_NO_DEVICE, "scratch: invalid device %d", device);
    if (static_cast<size_t>(device) >= g_scratch_ptr.size()) g_scratch_ptr.resize(device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e, "scratch: hipGetSymbolAddress");
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return SUPRA_OK;
}

}  // namespace

extern "C" int supra_bf_create(const supra_rx_geometry* g, supra_bf_ctx** out_ctx) {
    clear_error();
    if (!g || !out_ctx) return fail(SUPRA_ERR_INVALID_ARG, "create: null argument");
    *out_ctx = nullptr;
    if (!g->scanlines || !g->depths || !g->element_x || !g->element_y)
        return fail(SUPRA_ERR_INVALID_ARG, "create: geometry arrays must be non-null");
    if (g->num_rx_scanlines == 0 || g->num_depths == 0 || g->num_elements == 0)
        return fail(SUPRA_ERR_INVALID_ARG, "create: empty geometry");
    if (static_cast<uint64_t>(g->rx_scanline_layout[0]) * g->rx_scanline_layout[1] != g->num_rx_scanlines)
        return fail(SUPRA_ERR_INVALID_ARG, "create: rx_scanline_layout does not multiply to num_rx_scanlines");
    if (static_cast<uint64_t>(g->element_layout[0]) * g->element_layout[1] != g->num_elements)
        return fail(SUPRA_ERR_INVALID_ARG, "create: element_layout does not multiply to num_elements");
    if (g->num_rx_scanlines > (1u << 24) || g->num_depths > (1u << 24))
        return fail(SUPRA_ERR_TOO_LARGE, "create: geometry too large");

    supra_bf_ctx* ctx = new (std::nothrow) supra_bf_ctx();
    if (!ctx) return fail(SUPRA_ERR_INVALID_ARG, "create: out of host memory");
    ctx->n_rx = g->num_rx_scanlines;
    ctx->rx_layout[0] = g->rx_scanline_layout[0];
    ctx->rx_layout[1] = g->rx_scanline_layout[1];
    ctx->n_depths = g->num_depths;
    ctx->n_el = g->num_elements;
    ctx->el_layout[0] = g->element_layout[0];
    ctx->el_layout[1] = g->element_layout[1];
    ctx->is3d = g->rx_scanline_layout[0] > 1 && g->rx_scanline_layout[1] > 1;
    ctx->scan_mm.assign(g->scanlines, g->scanlines + g->num_rx_scanlines);
    ctx->depths_mm.assign(g->depths, g->depths + g->num_depths);
    ctx->ex_mm.assign(g->element_x, g->element_x + g->num_elements);
    ctx->ey_mm.assign(g->element_y, g->element_y + g->num_elements);
    ctx->geom_c = g->speed_of_sound_mm_per_s;
    ctx->depth_mm = g->depth;
    ctx->far_ok_mm = true;
    for (uint32_t i = 0; i < ctx->n_depths; i++)
        if (!(ctx->depths_mm[i] >= 0.0f) || (i > 0 && !(ctx->depths_mm[i] >= ctx->depths_mm[i - 1])))
            ctx->far_ok_mm = false;
    if (g->element_to_channel) {
        ctx->e2c.assign(g->element_to_channel, g->element_to_channel + g->num_elements);
        for (int32_t m : ctx->e2c) {
            if (m < SUPRA_ELEMENT_NOT_CONNECTED) {
                delete ctx;
                return fail(SUPRA_ERR_INVALID_ARG, "create: element_to_channel entries must be >= -1");
            }
            if (m >= 0) ctx->max_channel_in_map = std::max(ctx->max_channel_in_map, static_cast<uint32_t>(m));
        }
    }
    // Aperture corners must address existing elements (the reference reads
    // element tables unchecked).
    for (uint32_t i = 0; i < ctx->n_rx; i++) {
        for (int k = 0; k < 4; k++) {
            const supra_tx_params& t = ctx->scan_mm[i].tx[k];
            bool bad = ctx->is3d ? (t.last_active_element[0] > ctx->el_layout[0] ||
                                    t.last_active_element[1] > ctx->el_layout[1])
                                 : (t.last_active_element[0] > ctx->n_el);
            if (bad) {
                delete ctx;
                return fail(SUPRA_ERR_INVALID_ARG, "create: scanline %u transmit %d aperture exceeds the array", i, k);
            }
            if (t.last_active_element[0] > t.first_active_element[0])
                ctx->max_aperture = std::max<uint32_t>(ctx->max_aperture,
                                                       t.last_active_element[0] - t.first_active_element[0]);
        }
    }
    hipError_t e = hipGetDevice(&ctx->device);
    if (e != hipSuccess) {
        delete ctx;
        return fail(SUPRA_ERR_NO_DEVICE, "create: no HIP device: %s", hipGetErrorString(e));
    }
    if ((e = hipEventCreateWithFlags(&ctx->tables_used, hipEventDisableTiming)) != hipSuccess) {
        delete ctx;
        return hip_fail(e, "create: event");
    }
    if ((e = hipMalloc(&ctx->d_scans, ctx->n_rx * sizeof(ScanlineDT))) != hipSuccess ||
        (e = hipMalloc(&ctx->d_depths, ctx->n_depths * sizeof(float))) != hipSuccess ||
        (e = hipMalloc(&ctx->d_ex, ctx->n_el * sizeof(float))) != hipSuccess ||
        (e = hipMalloc(&ctx->d_ey, ctx->n_el * sizeof(float))) != hipSuccess) {
        delete ctx;
        return hip_fail(e, "create: hipMalloc");
    }
    if (!ctx->e2c.empty()) {
        if ((e = hipMalloc(&ctx->d_e2c, ctx->n_el * sizeof(int32_t))) != hipSuccess ||
            (e = hipMemcpy(ctx->d_e2c, ctx->e2c.data(), ctx->n_el * sizeof(int32_t), hipMemcpyHostToDevice)) !=
                hipSuccess) {
            delete ctx;
            return hip_fail(e, "create: element map upload");
        }
    }
    // The differentiable path's device geometry is set up here, so that its
    // allocation-free entry point allocates nothing on any call.
    int st = ctx->is3d ? SUPRA_OK : ensure_mm_geometry(ctx);
    if (st == SUPRA_OK) st = noalloc_scratch(ctx->device, &ctx->noalloc_scratch);
    if (st != SUPRA_OK) {
        delete ctx;
        return st;
    }
    *out_ctx = ctx;
    return SUPRA_OK;
}

extern "C" void supra_bf_destroy(supra_bf_ctx* ctx) { delete ctx; }

extern "C" int supra_das_run_frames(supra_bf_ctx* ctx, int32_t sample_beamformer, uint32_t num_frames,
                                    const void* rf, int32_t rf_dtype, uint32_t num_tx, uint32_t num_channels,
                                    uint32_t num_samples, double sampling_frequency_hz, double f_number,
                                    double speed_of_sound_mm_per_s, int32_t window_type, float window_parameter,
                                    int32_t interpolate_transmits, int32_t additional_offset, void* out,
                                    int32_t out_dtype, void* stream) {
    clear_error();
    if (!ctx || !rf || !out) return fail(SUPRA_ERR_INVALID_ARG, "das: null argument");
    if (num_frames == 0) return SUPRA_OK;
    if ((rf_dtype != SUPRA_INT16 && rf_dtype != SUPRA_FLOAT32) ||
        (out_dtype != SUPRA_INT16 && out_dtype != SUPRA_FLOAT32))
        return fail(SUPRA_ERR_UNSUPPORTED, "das: element types must be SUPRA_INT16 or SUPRA_FLOAT32");
    if (sample_beamformer < SUPRA_BF_DELAY_AND_SUM || sample_beamformer > SUPRA_BF_TEST_SIGNAL)
        return fail(SUPRA_ERR_UNSUPPORTED,
                    "das: unknown sample beamformer %d (DelayAndSum=0, DelayAndStdDev=1, "
                    "CoherenceFactorDelayAndSum=2, TestSignal=3)", sample_beamformer);
    if (window_type < SUPRA_WINDOW_RECTANGULAR || window_type > SUPRA_WINDOW_GAUSS)
        return fail(SUPRA_ERR_INVALID_ARG, "das: unknown window type %d", window_type);
    if (num_tx == 0 || num_channels == 0 || num_samples == 0)
        return fail(SUPRA_ERR_INVALID_ARG, "das: empty RF (num_tx=%u num_channels=%u num_samples=%u)", num_tx,
                    num_channels, num_samples);
    if (static_cast<uint64_t>(num_channels) * num_samples * (rf_dtype == SUPRA_INT16 ? 2 : 4) >= (1ull << 31))
        return fail(SUPRA_ERR_TOO_LARGE, "das: one transmit's RF block must be smaller than 2 GiB");
    if (!(sampling_frequency_hz > 0) || !(speed_of_sound_mm_per_s > 0) || !(f_number > 0) ||
        !std::isfinite(sampling_frequency_hz) || !std::isfinite(speed_of_sound_mm_per_s) || !std::isfinite(f_number))
        return fail(SUPRA_ERR_INVALID_ARG, "das: sampling frequency, speed of sound and f-number must be positive");
    if (!ctx->e2c.empty() && ctx->max_channel_in_map >= num_channels)
        return fail(SUPRA_ERR_INVALID_ARG, "das: element map addresses channel %u of %u", ctx->max_channel_in_map,
                    num_channels);
    // Transmit indices the kernels will read (RxBeamformerCuda.cu:299-303
    // leaves such pixels unwritten; here the call is rejected instead).
    for (uint32_t i = 0; i < ctx->n_rx; i++) {
        const supra_scanline_rx& s = ctx->scan_mm[i];
        int32_t sel, mask;
        float w[4];
        scanline_selection(s, &sel, &mask, w);
        for (int k = 0; k < 4; k++) {
            bool used = interpolate_transmits ? ((mask >> k) & 1) : (k == sel);
            if (used && s.tx[k].tx_scanline_idx >= num_tx)
                return fail(SUPRA_ERR_INVALID_ARG, "das: scanline %u uses transmit %u of %u", i,
                            static_cast<unsigned>(s.tx[k].tx_scanline_idx), num_tx);
        }
    }
    const uint64_t tile = das_depth_tile();
    const uint64_t blocks = ((ctx->n_depths + tile - 1) / tile) * ctx->n_rx * static_cast<uint64_t>(num_frames);
    if (blocks >= (1ull << 31)) return fail(SUPRA_ERR_TOO_LARGE, "das: too many frames for one launch");

    std::lock_guard<std::mutex> lock(ctx->mu);
    int st = check_device(ctx);
    if (st != SUPRA_OK) return st;
    hipStream_t hs = static_cast<hipStream_t>(stream);
    const double dt = 1.0 / sampling_frequency_hz;
    st = ensure_dt_tables(ctx, dt, speed_of_sound_mm_per_s, hs);
    if (st != SUPRA_OK) return st;

    DasArgs a;
    std::memset(&a, 0, sizeof(a));
    a.rf = rf;
    a.out = out;
    a.scans = ctx->d_scans;
    a.depths = ctx->d_depths;
    a.ex = ctx->d_ex;
    a.ey = ctx->d_ey;
    a.e2c = ctx->d_e2c;
    a.rf_frame_stride = static_cast<uint64_t>(num_tx) * num_channels * num_samples;
    a.out_frame_stride = static_cast<uint64_t>(ctx->n_rx) * ctx->n_depths;
    a.n_rx = ctx->n_rx;
    a.n_depths = ctx->n_depths;
    a.n_tx = num_tx;
    a.n_ch = num_channels;
    a.n_s = num_samples;
    a.n_el = ctx->n_el;
    a.layout_x = ctx->el_layout[0];
    a.layout_y = ctx->el_layout[1];
    a.n_frames = num_frames;
    a.depth_tiles = static_cast<uint32_t>((ctx->n_depths + tile - 1) / tile);
    a.F = static_cast<float>(f_number);
    a.dt = static_cast<float>(dt);
    a.c = static_cast<float>(speed_of_sound_mm_per_s);
    a.offset = static_cast<float>(additional_offset);
    a.far_ok = ctx->far_ok ? 1 : 0;
    build_window_lut(window_type, window_parameter, a.lut);

    hipError_t e;
    if (sample_beamformer == SUPRA_BF_DELAY_AND_STD_DEV || sample_beamformer == SUPRA_BF_COHERENCE_FACTOR)
        e = launch_das_variant(a, rf_dtype, out_dtype, ctx->is3d, sample_beamformer, interpolate_transmits != 0, hs);
    else
        e = launch_das(a, rf_dtype, out_dtype, ctx->is3d, sample_beamformer == SUPRA_BF_TEST_SIGNAL,
                       interpolate_transmits != 0, ctx->d_e2c != nullptr, hs);
    if (e != hipSuccess) return hip_fail(e, "das: launch");
    return tables_used_on(ctx, hs);
}

extern "C" int supra_das_run(supra_bf_ctx* ctx, int32_t sample_beamformer, const void* rf, int32_t rf_dtype,
                             uint32_t num_tx, uint32_t num_channels, uint32_t num_samples,
                             double sampling_frequency_hz, double f_number, double speed_of_sound_mm_per_s,
                             int32_t window_type, float window_parameter, int32_t interpolate_transmits,
                             int32_t additional_offset, void* out, int32_t out_dtype, void* stream) {
    return supra_das_run_frames(ctx, sample_beamformer, 1, rf, rf_dtype, num_tx, num_channels, num_samples,
                                sampling_frequency_hz, f_number, speed_of_sound_mm_per_s, window_type,
                                window_parameter, interpolate_transmits, additional_offset, out, out_dtype, stream);
}

namespace supra_amd {

// Checks of the pre-delay entry points (RawDelay::performDelay preconditions).
int check_rawdelay_call(const supra_bf_ctx* ctx, const char* what, int32_t rf_dtype, uint32_t num_tx,
                        uint32_t num_channels, uint32_t num_samples, double fs, double f_number, double c,
                        int32_t window_type, bool allow3d) {
    if (rf_dtype != SUPRA_INT16 && rf_dtype != SUPRA_FLOAT32)
        return fail(SUPRA_ERR_UNSUPPORTED, "%s: element types must be SUPRA_INT16 or SUPRA_FLOAT32", what);
    if (ctx->is3d && !allow3d) return fail(SUPRA_ERR_UNSUPPORTED, "%s: 2D geometries only", what);
    if (window_type < SUPRA_WINDOW_RECTANGULAR || window_type > SUPRA_WINDOW_GAUSS)
        return fail(SUPRA_ERR_INVALID_ARG, "%s: unknown window type %d", what, window_type);
    if (num_tx == 0 || num_channels == 0 || num_samples == 0) return fail(SUPRA_ERR_INVALID_ARG, "%s: empty RF", what);
    if (static_cast<uint64_t>(num_channels) * num_samples * (rf_dtype == SUPRA_INT16 ? 2 : 4) >= (1ull << 31))
        return fail(SUPRA_ERR_TOO_LARGE, "%s: one transmit's RF block must be smaller than 2 GiB", what);
    if (!(fs > 0) || !(c > 0) || !(f_number > 0) || !std::isfinite(fs) || !std::isfinite(c) || !std::isfinite(f_number))
        return fail(SUPRA_ERR_INVALID_ARG, "%s: sampling frequency, speed of sound and f-number must be positive",
                    what);
    if (!ctx->e2c.empty())
        return fail(SUPRA_ERR_UNSUPPORTED, "%s: element-to-channel maps are not used by the pre-delay", what);
    for (uint32_t i = 0; i < ctx->n_rx; i++) {
        const supra_scanline_rx& s = ctx->scan_mm[i];
        int32_t sel, mask;
        float w[4];
        scanline_selection(s, &sel, &mask, w);
        if (sel < 0) continue;
        const supra_tx_params& t = s.tx[sel];
        if (t.tx_scanline_idx >= num_tx)
            return fail(SUPRA_ERR_INVALID_ARG, "%s: scanline %u uses transmit %u of %u", what, i,
                        static_cast<unsigned>(t.tx_scanline_idx), num_tx);
        if (ctx->is3d) {
            // slot (ex - first.x) + (ey - first.y) * layout.x (RawDelay.cu:207-209)
            if (t.last_active_element[0] > t.first_active_element[0] &&
                t.last_active_element[1] > t.first_active_element[1] &&
                static_cast<uint64_t>(t.last_active_element[0] - t.first_active_element[0] - 1) +
                        static_cast<uint64_t>(t.last_active_element[1] - t.first_active_element[1] - 1) *
                            ctx->el_layout[0] >= num_channels)
                return fail(SUPRA_ERR_INVALID_ARG, "%s: scanline %u aperture exceeds the %u output channels", what, i,
                            num_channels);
        } else if (t.last_active_element[0] > t.first_active_element[0] &&
                   static_cast<uint32_t>(t.last_active_element[0] - t.first_active_element[0]) > num_channels) {
            return fail(SUPRA_ERR_INVALID_ARG, "%s: scanline %u aperture exceeds the %u output channels", what, i,
                        num_channels);
        }
    }
    return SUPRA_OK;
}

// DasArgs of a one-frame 2D pre-delay (tables must be current).
DasArgs rawdelay_args(const supra_bf_ctx* ctx, const void* rf, uint32_t num_tx, uint32_t num_channels,
                      uint32_t num_samples, double dt, double f_number, double c, int32_t window_type,
                      float window_parameter) {
    DasArgs a;
    std::memset(&a, 0, sizeof(a));
    a.rf = rf;
    a.scans = ctx->d_scans;
    a.depths = ctx->d_depths;
    a.ex = ctx->d_ex;
    a.ey = ctx->d_ey;
    a.rf_frame_stride = static_cast<uint64_t>(num_tx) * num_channels * num_samples;
    a.out_frame_stride = static_cast<uint64_t>(ctx->n_rx) * num_channels * ctx->n_depths;
    a.n_rx = ctx->n_rx;
    a.n_depths = ctx->n_depths;
    a.n_tx = num_tx;
    a.n_ch = num_channels;
    a.n_s = num_samples;
    a.n_el = ctx->n_el;
    a.layout_x = ctx->el_layout[0];
    a.layout_y = ctx->el_layout[1];
    a.n_frames = 1;
    a.depth_tiles = (ctx->n_depths + das_depth_tile() - 1) / das_depth_tile();
    a.F = static_cast<float>(f_number);
    a.dt = static_cast<float>(dt);
    a.c = static_cast<float>(c);
    a.far_ok = 0;
    build_window_lut(window_type, window_parameter, a.lut);
    return a;
}

}  // namespace

extern "C" int supra_rawdelay_run(supra_bf_ctx* ctx, const void* rf, int32_t rf_dtype, uint32_t num_tx,
                                  uint32_t num_channels, uint32_t num_samples, double sampling_frequency_hz,
                                  double f_number, double speed_of_sound_mm_per_s, int32_t window_type,
                                  float window_parameter, void* out, int32_t out_dtype,
                                  double* out_sampling_frequency_hz, void* stream) {
    clear_error();
    if (!ctx || !rf || !out) return fail(SUPRA_ERR_INVALID_ARG, "rawdelay: null argument");
    if (out_dtype != SUPRA_INT16 && out_dtype != SUPRA_FLOAT32)
        return fail(SUPRA_ERR_UNSUPPORTED, "rawdelay: element types must be SUPRA_INT16 or SUPRA_FLOAT32");
    int st = check_rawdelay_call(ctx, "rawdelay", rf_dtype, num_tx, num_channels, num_samples, sampling_frequency_hz,
                                 f_number, speed_of_sound_mm_per_s, window_type, true);
    if (st != SUPRA_OK) return st;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if ((st = check_device(ctx)) != SUPRA_OK) return st;
    hipStream_t hs = static_cast<hipStream_t>(stream);
    const double dt = 1.0 / sampling_frequency_hz;
    if ((st = ensure_dt_tables(ctx, dt, speed_of_sound_mm_per_s, hs)) != SUPRA_OK) return st;
    DasArgs a = rawdelay_args(ctx, rf, num_tx, num_channels, num_samples, dt, f_number, speed_of_sound_mm_per_s,
                              window_type, window_parameter);
    a.out = out;
    hipError_t e = launch_rawdelay(a, rf_dtype, out_dtype, ctx->is3d, hs);
    if (e != hipSuccess) return hip_fail(e, "rawdelay: launch");
    if ((st = tables_used_on(ctx, hs)) != SUPRA_OK) return st;
    if (out_sampling_frequency_hz) {
        // RawDelay.cu:48, :642: depth of the last sample / c / N_d.
        const double depth = ctx->depths_mm.back();
        *out_sampling_frequency_hz = 1 / (depth / speed_of_sound_mm_per_s / ctx->n_depths);
    }
    return SUPRA_OK;
}

extern "C" int supra_mv_das_run(supra_bf_ctx* ctx, const void* rf, int32_t rf_dtype, uint32_t num_tx,
                                uint32_t num_channels, uint32_t num_samples, double sampling_frequency_hz,
                                double f_number, double speed_of_sound_mm_per_s, int32_t window_type,
                                float window_parameter, uint32_t sub_array_size, uint32_t temporal_smoothing,
                                double sub_array_scaling_power, int32_t compute_means, double loading_factor,
                                void* out, int32_t out_dtype, void* stream) {
    clear_error();
    if (!ctx || !rf || !out) return fail(SUPRA_ERR_INVALID_ARG, "mv: null argument");
    if (out_dtype != SUPRA_INT16 && out_dtype != SUPRA_FLOAT32)
        return fail(SUPRA_ERR_UNSUPPORTED, "mv: element types must be SUPRA_INT16 or SUPRA_FLOAT32");
    int st = check_rawdelay_call(ctx, "mv", rf_dtype, num_tx, num_channels, num_samples, sampling_frequency_hz,
                                 f_number, speed_of_sound_mm_per_s, window_type, false);
    if (st != SUPRA_OK) return st;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if ((st = check_device(ctx)) != SUPRA_OK) return st;
    hipStream_t hs = static_cast<hipStream_t>(stream);
    const double dt = 1.0 / sampling_frequency_hz;
    if ((st = ensure_dt_tables(ctx, dt, speed_of_sound_mm_per_s, hs)) != SUPRA_OK) return st;
    const DasArgs g = rawdelay_args(ctx, rf, num_tx, num_channels, num_samples, dt, f_number,
                                    speed_of_sound_mm_per_s, window_type, window_parameter);
    st = launch_mv_fused(g, rf_dtype, sub_array_size, temporal_smoothing, sub_array_scaling_power, compute_means,
                         loading_factor, out, out_dtype, hs);
    if (st != SUPRA_OK) return st;
    return tables_used_on(ctx, hs);
}

// ---- differentiable 2D DAS -----------------------------------------------------

namespace {

int ensure_capacity(void** p, size_t* cap, size_t bytes, const char* what) {
    if (bytes <= *cap) return SUPRA_OK;
    if (*p) {
        hipError_t e = hipFree(*p);
        *p = nullptr;
        *cap = 0;
        if (e != hipSuccess) return hip_fail(e, what);
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return hip_fail(e, what);
    }
    *cap = bytes;
    return SUPRA_OK;
}

size_t align256(size_t n) { return (n + 255) & ~static_cast<size_t>(255); }

// The millimetre geometry and the c-independent fields of the scanline
// records (transmit selection, apertures, directions) on the device.
int ensure_mm_geometry(supra_bf_ctx* ctx) {
    if (ctx->d_scan_mm) return SUPRA_OK;
    std::vector<ScanlineDT> fixed(ctx->n_rx);
    for (uint32_t i = 0; i < ctx->n_rx; i++) {
        const supra_scanline_rx& s = ctx->scan_mm[i];
        ScanlineDT& o = fixed[i];
        std::memset(&o, 0, sizeof(o));
        o.dir_x = static_cast<float>(s.direction[0]);
        o.dir_y = static_cast<float>(s.direction[1]);
        o.dir_z = static_cast<float>(s.direction[2]);
        scanline_selection(s, &o.sel, &o.active_mask, o.txw);
        for (int k = 0; k < 4; k++) {
            o.tx[k].first_x = s.tx[k].first_active_element[0];
            o.tx[k].first_y = s.tx[k].first_active_element[1];
            o.tx[k].last_x = s.tx[k].last_active_element[0];
            o.tx[k].last_y = s.tx[k].last_active_element[1];
            o.tx[k].tx_idx = s.tx[k].tx_scanline_idx;
        }
    }
    hipError_t e;
    if ((e = hipMalloc(&ctx->d_scan_mm, ctx->n_rx * sizeof(supra_scanline_rx))) != hipSuccess ||
        (e = hipMalloc(&ctx->d_scan_fixed, ctx->n_rx * sizeof(ScanlineDT))) != hipSuccess ||
        (e = hipMalloc(&ctx->d_depths_mm, ctx->n_depths * sizeof(float))) != hipSuccess ||
        (e = hipMalloc(&ctx->d_ex_mm, ctx->n_el * sizeof(float))) != hipSuccess ||
        (e = hipMemcpy(ctx->d_scan_mm, ctx->scan_mm.data(), ctx->n_rx * sizeof(supra_scanline_rx),
                       hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(ctx->d_scan_fixed, fixed.data(), ctx->n_rx * sizeof(ScanlineDT), hipMemcpyHostToDevice)) !=
            hipSuccess ||
        (e = hipMemcpy(ctx->d_depths_mm, ctx->depths_mm.data(), ctx->n_depths * sizeof(float),
                       hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(ctx->d_ex_mm, ctx->ex_mm.data(), ctx->n_el * sizeof(float), hipMemcpyHostToDevice)) !=
            hipSuccess) {
        (void)hipFree(ctx->d_scan_mm);
        ctx->d_scan_mm = nullptr;
        return hip_fail(e, "differentiable DAS: geometry upload");
    }
    return SUPRA_OK;
}

struct FrameTables {
    ScanlineDT* scans;
    float* depths;
    float* ex;
    float* c;
};

// Bytes of the per-frame tables of `frames` frames (256-byte aligned parts).
size_t frame_table_bytes(const supra_bf_ctx* ctx, uint32_t frames) {
    return align256(static_cast<size_t>(frames) * ctx->n_rx * sizeof(ScanlineDT)) +
           align256(static_cast<size_t>(frames) * ctx->n_depths * sizeof(float)) +
           align256(static_cast<size_t>(frames) * ctx->n_el * sizeof(float)) +
           align256(static_cast<size_t>(frames) * sizeof(float));
}

// dt-space tables of every frame for the device array c (no host round trip),
// built at `base` (frame_table_bytes(ctx, frames) bytes of device memory).
int build_frame_tables_at(supra_bf_ctx* ctx, void* base_ptr, uint32_t frames, const double* c_dev, double dt,
                          hipStream_t stream, FrameTables* out) {
    const size_t b_scans = align256(static_cast<size_t>(frames) * ctx->n_rx * sizeof(ScanlineDT));
    const size_t b_depths = align256(static_cast<size_t>(frames) * ctx->n_depths * sizeof(float));
    const size_t b_ex = align256(static_cast<size_t>(frames) * ctx->n_el * sizeof(float));
    char* base = static_cast<char*>(base_ptr);
    out->scans = reinterpret_cast<ScanlineDT*>(base);
    out->depths = reinterpret_cast<float*>(base + b_scans);
    out->ex = reinterpret_cast<float*>(base + b_scans + b_depths);
    out->c = reinterpret_cast<float*>(base + b_scans + b_depths + b_ex);
    FrameTablesArgs a;
    std::memset(&a, 0, sizeof(a));
    a.scan_mm = ctx->d_scan_mm;
    a.scan_fixed = ctx->d_scan_fixed;
    a.depths_mm = ctx->d_depths_mm;
    a.ex_mm = ctx->d_ex_mm;
    a.e2c = ctx->d_e2c;
    a.c = c_dev;
    a.scans = out->scans;
    a.depths = out->depths;
    a.ex = out->ex;
    a.frame_c = out->c;
    a.n_rx = ctx->n_rx;
    a.n_d = ctx->n_depths;
    a.n_el = ctx->n_el;
    a.n_frames = frames;
    a.dt = dt;
    hipError_t e = launch_frame_tables(a, stream);
    if (e != hipSuccess) return hip_fail(e, "differentiable DAS: table launch");
    return SUPRA_OK;
}

// The same in the context's own table memory, grown on demand.
int build_frame_tables(supra_bf_ctx* ctx, uint32_t frames, const double* c_dev, double dt, hipStream_t stream,
                       FrameTables* out) {
    int st = ensure_capacity(&ctx->d_frame_tables, &ctx->frame_tables_cap, frame_table_bytes(ctx, frames),
                             "differentiable DAS: table memory");
    if (st != SUPRA_OK) return st;
    return build_frame_tables_at(ctx, ctx->d_frame_tables, frames, c_dev, dt, stream, out);
}

// Which (transmit, channel) RF lines the adjoint writes and how often: with
// at most one writer per line the lines are stored without atomics and the
// never-written lines are zeroed from a list.
int ensure_line_writers(supra_bf_ctx* ctx, int interp, uint32_t num_tx, uint32_t num_channels) {
    if (ctx->lines_key_interp == interp && ctx->lines_key_tx == num_tx && ctx->lines_key_ch == num_channels)
        return SUPRA_OK;
    std::vector<uint8_t> cnt(static_cast<size_t>(num_tx) * num_channels, 0);
    for (uint32_t i = 0; i < ctx->n_rx; i++) {
        const supra_scanline_rx& s = ctx->scan_mm[i];
        int32_t sel, mask;
        float w[4];
        scanline_selection(s, &sel, &mask, w);
        for (int k = 0; k < 4; k++) {
            const bool used = interp ? ((mask >> k) & 1) != 0 : (k == sel);
            if (!used) continue;
            const supra_tx_params& t = s.tx[k];
            for (uint32_t e = t.first_active_element[0]; e < t.last_active_element[0]; e++) {
                uint32_t ch;
                if (!ctx->e2c.empty()) {
                    if (ctx->e2c[e] < 0) continue;
                    ch = static_cast<uint32_t>(ctx->e2c[e]);
                } else {
                    ch = e % num_channels;
                }
                uint8_t& c = cnt[static_cast<size_t>(t.tx_scanline_idx) * num_channels + ch];
                c = static_cast<uint8_t>(std::min(2, c + 1));
            }
        }
    }
    std::vector<uint32_t> zero;
    bool exclusive = true;
    for (size_t i = 0; i < cnt.size(); i++) {
        if (cnt[i] > 1) exclusive = false;
        if (cnt[i] == 0) zero.push_back(static_cast<uint32_t>(i));
    }
    (void)hipFree(ctx->d_zero_lines);
    ctx->d_zero_lines = nullptr;
    ctx->n_zero_lines = 0;
    ctx->lines_key_interp = -1;
    if (exclusive && !zero.empty()) {
        hipError_t e;
        if ((e = hipMalloc(&ctx->d_zero_lines, zero.size() * sizeof(uint32_t))) != hipSuccess ||
            (e = hipMemcpy(ctx->d_zero_lines, zero.data(), zero.size() * sizeof(uint32_t), hipMemcpyHostToDevice)) !=
                hipSuccess)
            return hip_fail(e, "adjoint: line list upload");
        ctx->n_zero_lines = static_cast<uint32_t>(zero.size());
    }
    ctx->lines_exclusive = exclusive;
    ctx->lines_key_interp = interp;
    ctx->lines_key_tx = num_tx;
    ctx->lines_key_ch = num_channels;
    return SUPRA_OK;
}

// Argument checks shared by the differentiable entry points.
int check_diff_call(const supra_bf_ctx* ctx, const char* what, uint32_t num_frames, uint32_t num_tx,
                    uint32_t num_channels, uint32_t num_samples, double fs, double f_number, int32_t window_type,
                    int32_t interpolate_transmits) {
    if (ctx->is3d) return fail(SUPRA_ERR_UNSUPPORTED, "%s: 2D geometries only", what);
    if (num_frames > 65535) return fail(SUPRA_ERR_TOO_LARGE, "%s: at most 65535 frames per call", what);
    if (window_type < SUPRA_WINDOW_RECTANGULAR || window_type > SUPRA_WINDOW_GAUSS)
        return fail(SUPRA_ERR_INVALID_ARG, "%s: unknown window type %d", what, window_type);
    if (num_tx == 0 || num_channels == 0 || num_samples == 0)
        return fail(SUPRA_ERR_INVALID_ARG, "%s: empty RF (num_tx=%u num_channels=%u num_samples=%u)", what, num_tx,
                    num_channels, num_samples);
    if (static_cast<uint64_t>(num_channels) * num_samples * 4 >= (1ull << 31))
        return fail(SUPRA_ERR_TOO_LARGE, "%s: one transmit's RF block must be smaller than 2 GiB", what);
    if (!(fs > 0) || !(f_number > 0) || !std::isfinite(fs) || !std::isfinite(f_number))
        return fail(SUPRA_ERR_INVALID_ARG, "%s: sampling frequency and f-number must be positive", what);
    if (!ctx->e2c.empty() && ctx->max_channel_in_map >= num_channels)
        return fail(SUPRA_ERR_INVALID_ARG, "%s: element map addresses channel %u of %u", what,
                    ctx->max_channel_in_map, num_channels);
    for (uint32_t i = 0; i < ctx->n_rx; i++) {
        const supra_scanline_rx& s = ctx->scan_mm[i];
        int32_t sel, mask;
        float w[4];
        scanline_selection(s, &sel, &mask, w);
        for (int k = 0; k < 4; k++) {
            bool used = interpolate_transmits ? ((mask >> k) & 1) : (k == sel);
            if (used && s.tx[k].tx_scanline_idx >= num_tx)
                return fail(SUPRA_ERR_INVALID_ARG, "%s: scanline %u uses transmit %u of %u", what, i,
                            static_cast<unsigned>(s.tx[k].tx_scanline_idx), num_tx);
        }
    }
    return SUPRA_OK;
}

}  // namespace

namespace {

// supra_das_forward_c with its per-frame tables either in the context's own
// memory (noalloc == false) or in this device's file-scope scratch array.
int das_forward_c(supra_bf_ctx* ctx, bool noalloc, uint32_t num_frames, const float* rf, uint32_t num_tx,
                  uint32_t num_channels, uint32_t num_samples, double sampling_frequency_hz, double f_number,
                  const double* speed_of_sound_mm_per_s, int32_t window_type, float window_parameter,
                  int32_t interpolate_transmits, int32_t additional_offset, float* out, float* norm_out,
                  void* stream) {
    const char* what = noalloc ? "das forward (no allocation)" : "das forward";
    if (!ctx || !rf || !out || !speed_of_sound_mm_per_s) return fail(SUPRA_ERR_INVALID_ARG, "%s: null argument", what);
    if (num_frames == 0) return SUPRA_OK;
    int st = check_diff_call(ctx, what, num_frames, num_tx, num_channels, num_samples, sampling_frequency_hz,
                             f_number, window_type, interpolate_transmits);
    if (st != SUPRA_OK) return st;
    const uint64_t tile = das_depth_tile();
    const uint64_t blocks = ((ctx->n_depths + tile - 1) / tile) * ctx->n_rx * static_cast<uint64_t>(num_frames);
    if (blocks >= (1ull << 31)) return fail(SUPRA_ERR_TOO_LARGE, "%s: too many frames for one launch", what);
    if (noalloc && frame_table_bytes(ctx, num_frames) > SUPRA_NOALLOC_SCRATCH_BYTES)
        return fail(SUPRA_ERR_TOO_LARGE, "%s: the tables of %u frames need %zu bytes, more than the %u-byte scratch",
                    what, num_frames, frame_table_bytes(ctx, num_frames), SUPRA_NOALLOC_SCRATCH_BYTES);

    std::lock_guard<std::mutex> lock(ctx->mu);
    if ((st = check_device(ctx)) != SUPRA_OK) return st;
    hipStream_t hs = static_cast<hipStream_t>(stream);
    const double dt = 1.0 / sampling_frequency_hz;
    FrameTables ft;
    if (noalloc) {
        if (!ctx->d_scan_mm || !ctx->noalloc_scratch)
            return fail(SUPRA_ERR_INVALID_ARG, "%s: context has no device geometry", what);
        st = build_frame_tables_at(ctx, ctx->noalloc_scratch, num_frames, speed_of_sound_mm_per_s, dt, hs, &ft);
    } else {
        if ((st = ensure_mm_geometry(ctx)) != SUPRA_OK) return st;
        st = build_frame_tables(ctx, num_frames, speed_of_sound_mm_per_s, dt, hs, &ft);
    }
    if (st != SUPRA_OK) return st;

    DasArgs a;
    std::memset(&a, 0, sizeof(a));
    a.rf = rf;
    a.out = out;
    a.scans = ft.scans;
    a.depths = ft.depths;
    a.ex = ft.ex;
    a.ey = ft.ex;   // unused in 2D
    a.e2c = ctx->d_e2c;
    a.rf_frame_stride = static_cast<uint64_t>(num_tx) * num_channels * num_samples;
    a.out_frame_stride = static_cast<uint64_t>(ctx->n_rx) * ctx->n_depths;
    a.n_rx = ctx->n_rx;
    a.n_depths = ctx->n_depths;
    a.n_tx = num_tx;
    a.n_ch = num_channels;
    a.n_s = num_samples;
    a.n_el = ctx->n_el;
    a.layout_x = ctx->el_layout[0];
    a.layout_y = ctx->el_layout[1];
    a.n_frames = num_frames;
    a.depth_tiles = static_cast<uint32_t>((ctx->n_depths + tile - 1) / tile);
    a.F = static_cast<float>(f_number);
    a.dt = static_cast<float>(dt);
    a.c = 0.0f;
    a.offset = static_cast<float>(additional_offset);
    a.far_ok = ctx->far_ok_mm ? 1 : 0;
    a.scan_frame_stride = ctx->n_rx;
    a.depth_frame_stride = ctx->n_depths;
    a.el_frame_stride = ctx->n_el;
    a.frame_c = ft.c;
    a.norm_out = norm_out;
    build_window_lut(window_type, window_parameter, a.lut);
    hipError_t e = launch_das(a, SUPRA_FLOAT32, SUPRA_FLOAT32, false, false, interpolate_transmits != 0,
                              ctx->d_e2c != nullptr, hs);
    if (e != hipSuccess) return hip_fail(e, "das forward: launch");
    return SUPRA_OK;
}

}  // namespace

extern "C" int supra_das_forward_c(supra_bf_ctx* ctx, uint32_t num_frames, const float* rf, uint32_t num_tx,
                                   uint32_t num_channels, uint32_t num_samples, double sampling_frequency_hz,
                                   double f_number, const double* speed_of_sound_mm_per_s, int32_t window_type,
                                   float window_parameter, int32_t interpolate_transmits, int32_t additional_offset,
                                   float* out, float* norm_out, void* stream) {
    clear_error();
    return das_forward_c(ctx, false, num_frames, rf, num_tx, num_channels, num_samples, sampling_frequency_hz,
                         f_number, speed_of_sound_mm_per_s, window_type, window_parameter, interpolate_transmits,
                         additional_offset, out, norm_out, stream);
}

extern "C" int supra_das_forward_c_noalloc(supra_bf_ctx* ctx, uint32_t num_frames, const float* rf,
                                           uint32_t num_tx, uint32_t num_channels, uint32_t num_samples,
                                           double sampling_frequency_hz, double f_number,
                                           const double* speed_of_sound_mm_per_s, int32_t window_type,
                                           float window_parameter, int32_t interpolate_transmits,
                                           int32_t additional_offset, float* out, float* norm_out, void* stream) {
    clear_error();
    return das_forward_c(ctx, true, num_frames, rf, num_tx, num_channels, num_samples, sampling_frequency_hz,
                         f_number, speed_of_sound_mm_per_s, window_type, window_parameter, interpolate_transmits,
                         additional_offset, out, norm_out, stream);
}

extern "C" int supra_das_adjoint(supra_bf_ctx* ctx, uint32_t num_frames, const float* rf, uint32_t num_tx,
                                 uint32_t num_channels, uint32_t num_samples, double sampling_frequency_hz,
                                 double f_number, const double* speed_of_sound_mm_per_s, int32_t window_type,
                                 float window_parameter, int32_t interpolate_transmits, int32_t additional_offset,
                                 const float* norm, const float* grad_image, float* grad_rf, double* grad_c,
                                 void* stream) {
    clear_error();
    if (!ctx || !rf || !speed_of_sound_mm_per_s || !norm || !grad_image)
        return fail(SUPRA_ERR_INVALID_ARG, "das adjoint: null argument");
    if (num_frames == 0 || (!grad_rf && !grad_c)) return SUPRA_OK;
    int st = check_diff_call(ctx, "das adjoint", num_frames, num_tx, num_channels, num_samples,
                             sampling_frequency_hz, f_number, window_type, interpolate_transmits);
    if (st != SUPRA_OK) return st;
    if (ctx->n_rx > 65535u * 64u || ctx->n_depths > 65535u * 64u)
        return fail(SUPRA_ERR_TOO_LARGE, "das adjoint: image too large");

    std::lock_guard<std::mutex> lock(ctx->mu);
    if ((st = check_device(ctx)) != SUPRA_OK) return st;
    // Waves (aperture elements) per workgroup: each accumulates its RF line in
    // a fixed-size LDS ring, so any number of samples per channel fits.
    hipError_t e = hipSuccess;
    const uint32_t waves = 4;
    const uint32_t slots = interpolate_transmits ? 4u : 1u;
    const uint32_t egroups = std::max(1u, (ctx->max_aperture + waves - 1) / waves);
    const uint64_t blocks = static_cast<uint64_t>(num_frames) * ctx->n_rx * slots * egroups;
    if (blocks >= (1ull << 31)) return fail(SUPRA_ERR_TOO_LARGE, "das adjoint: too many frames for one launch");

    if ((st = ensure_mm_geometry(ctx)) != SUPRA_OK) return st;
    if ((st = ensure_line_writers(ctx, interpolate_transmits != 0, num_tx, num_channels)) != SUPRA_OK) return st;
    hipStream_t hs = static_cast<hipStream_t>(stream);
    const double dt = 1.0 / sampling_frequency_hz;
    FrameTables ft;
    if ((st = build_frame_tables(ctx, num_frames, speed_of_sound_mm_per_s, dt, hs, &ft)) != SUPRA_OK) return st;
    // beta, then the per-frame depth table (4 floats per depth), in one buffer
    const size_t n_beta = static_cast<size_t>(num_frames) * slots * ctx->n_rx * ctx->n_depths;
    const size_t b_beta = align256(n_beta * sizeof(float));
    const size_t b_dtab = static_cast<size_t>(num_frames) * ctx->n_depths * 4 * sizeof(float);
    if ((st = ensure_capacity(&ctx->d_beta, &ctx->beta_cap, b_beta + b_dtab, "das adjoint: scratch")) != SUPRA_OK)
        return st;
    float* dtab = reinterpret_cast<float*>(static_cast<char*>(ctx->d_beta) + b_beta);
    const size_t per_frame = static_cast<size_t>(ctx->n_rx) * slots * egroups * waves;
    if (grad_c && (st = ensure_capacity(&ctx->d_dc_part, &ctx->dc_part_cap, per_frame * num_frames * sizeof(double),
                                        "das adjoint: scratch")) != SUPRA_OK)
        return st;

    const uint64_t rf_frame = static_cast<uint64_t>(num_tx) * num_channels * num_samples;
    if (grad_rf) {
        if (!ctx->lines_exclusive) e = launch_fill_zero(grad_rf, rf_frame * num_frames, hs);
        else e = launch_zero_lines(grad_rf, ctx->d_zero_lines, ctx->n_zero_lines, num_samples, rf_frame, num_frames,
                                   hs);
        if (e != hipSuccess) return hip_fail(e, "das adjoint: zero launch");
    }
    e = launch_beta(grad_image, norm, static_cast<float*>(ctx->d_beta), ctx->n_rx, ctx->n_depths, slots, num_frames,
                    hs);
    if (e != hipSuccess) return hip_fail(e, "das adjoint: beta launch");
    e = launch_depth_table(ft.depths, ft.c, static_cast<float>(dt), static_cast<float>(f_number), ctx->n_depths,
                           num_frames, dtab, hs);
    if (e != hipSuccess) return hip_fail(e, "das adjoint: depth table launch");

    AdjointArgs a;
    std::memset(&a, 0, sizeof(a));
    a.rf = rf;
    a.beta = static_cast<const float*>(ctx->d_beta);
    a.scans = ft.scans;
    a.dtab = dtab;
    a.ex = ft.ex;
    a.frame_c = ft.c;
    a.e2c = ctx->d_e2c;
    a.grad_rf = grad_rf;
    a.dc_part = grad_c ? static_cast<double*>(ctx->d_dc_part) : nullptr;
    a.rf_frame_stride = rf_frame;
    a.scan_frame_stride = ctx->n_rx;
    a.depth_frame_stride = ctx->n_depths;
    a.el_frame_stride = ctx->n_el;
    a.n_rx = ctx->n_rx;
    a.n_depths = ctx->n_depths;
    a.n_ch = num_channels;
    a.n_s = num_samples;
    a.n_slots = slots;
    a.egroups = egroups;
    a.waves = waves;
    a.F = static_cast<float>(f_number);
    a.dt = static_cast<float>(dt);
    a.offset = static_cast<float>(additional_offset);
    a.exclusive = ctx->lines_exclusive ? 1 : 0;
    a.interp_tx = interpolate_transmits ? 1 : 0;
    build_window_lut(window_type, window_parameter, a.lut);
    if ((e = launch_das_adjoint(a, num_frames, hs)) != hipSuccess) return hip_fail(e, "das adjoint: launch");
    if (grad_c) {
        e = launch_dc_reduce(static_cast<const double*>(ctx->d_dc_part), static_cast<uint32_t>(per_frame),
                             speed_of_sound_mm_per_s, grad_c, num_frames, hs);
        if (e != hipSuccess) return hip_fail(e, "das adjoint: reduce launch");
    }
    return SUPRA_OK;
}
