// Batched beamforming of many acquisitions of uneven size in one graph
// launch (no reference counterpart: SUPRA processes one USRawData per node
// call, BeamformingNode.cpp:128-177, and one RawDelay + MV chain per frame).
//
// A batch is planned once: every acquisition becomes a frame of work items
// (scanline x tile of output depths) in one of three queues, by the register
// and LDS footprint of its kernel: delay-and-sum, minimum variance with
// L <= 32 and with L <= 64 (the pre-delay fused in).  Each queue is served by
// a persistent work-queue kernel sized to the GPU; workgroups take items in
// frame order, so frames complete in submission order within a queue, and
// the last item of a frame stamps the frame's completion time from the device
// wall clock.  The queues run concurrently (a fork/join of three streams) and
// the whole launch — reset of the queue counters, the three kernels — is
// captured once in a hipGraph; supra_batch_launch replays it.
//
// Footprints differ by an order of magnitude (DAS: 70 VGPRs, 10 KB LDS; MV:
// 256 VGPRs and up to the whole LDS), so one kernel for all items would run
// the delay-and-sum items at the minimum-variance occupancy; the queues keep
// each class at its own.
//
// The autotuner times each queue alone on the real batch for a few
// candidates (DAS: persistent workgroups per CU; MV: depth tile and LDS
// budget, i.e. one or two workgroups per CU) and keeps the fastest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "bf_ctx.hpp"
#include "bf_internal.hpp"
#include "das_launch.hpp"
#include "mv_launch.hpp"
#include "status.hpp"
#include "supra_bf.h"

using namespace supra_amd;

namespace {

constexpr uint32_t kQueues = 3;            // DAS, MV (L <= 32), MV (L <= 64)
constexpr uint32_t kMaxFramesPerQueue = 256;
constexpr uint32_t kMaxItemsPerFrame = 1u << 24;

__global__ void batch_reset_kernel(uint32_t* heads, uint32_t* done, uint64_t* stamp, uint32_t n_frames) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kQueues) heads[i] = 0;
    if (i < n_frames) {
        done[i] = 0;
        stamp[i] = 0;
    }
    if (i == 0) stamp[n_frames] = wall_clock64();   // launch start
}

struct FrameDesc {
    supra_batch_item item;
    uint32_t queue;            // 0 DAS, 1 MV32, 2 MV64
    double dt;
    size_t table_offset;       // bytes into the table block
};

}  // namespace

struct supra_batch {
    int device = 0;
    uint32_t n_frames = 0;
    int cus = 0;
    double clock_khz = 0;
    std::vector<FrameDesc> frames;
    // device memory
    void* tables = nullptr;            // per-frame dt-space tables
    size_t tables_bytes = 0;
    uint32_t* sync = nullptr;          // heads[kQueues] + done[n_frames] (uint32) + stamps (uint64)
    uint64_t* stamp = nullptr;
    uint32_t* heads = nullptr;
    uint32_t* done = nullptr;
    DasBatchFrame* das_frames = nullptr;
    uint32_t* das_items = nullptr;
    uint32_t das_n_items = 0, das_n_frames = 0;
    MvBatchFrame* mv_frames[2] = {nullptr, nullptr};
    uint32_t* mv_items[2] = {nullptr, nullptr};
    uint32_t mv_n_items[2] = {0, 0}, mv_n_frames[2] = {0, 0};
    size_t mv_lds[2] = {0, 0};
    // chosen configuration
    uint32_t das_wg_per_cu = 4;
    uint32_t mv_wg_per_cu[2] = {1, 1};
    size_t mv_budget[2] = {0, 0};
    std::string config;
    // graph
    hipStream_t streams[kQueues] = {nullptr, nullptr, nullptr};
    hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t last_stream = nullptr;

    ~supra_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        for (auto& e : ev_join)
            if (e) (void)hipEventDestroy(e);
        for (auto& s : streams)
            if (s) (void)hipStreamDestroy(s);
        (void)hipFree(tables);
        (void)hipFree(sync);
        (void)hipFree(das_frames);
        (void)hipFree(das_items);
        for (int q = 0; q < 2; q++) {
            (void)hipFree(mv_frames[q]);
            (void)hipFree(mv_items[q]);
        }
    }
};

namespace {

size_t align256(size_t n) { return (n + 255) & ~static_cast<size_t>(255); }

size_t table_bytes(const supra_bf_ctx* ctx) {
    return align256(ctx->n_rx * sizeof(ScanlineDT)) + align256(ctx->n_depths * sizeof(float)) +
           align256(ctx->n_el * sizeof(float));
}

// Checks of one batch entry (the preconditions of supra_das_run and
// supra_mv_das_run for the supported combinations).
int check_item(const supra_batch_item& it, uint32_t index) {
    if (!it.ctx || !it.rf || !it.out) return fail(SUPRA_ERR_INVALID_ARG, "batch item %u: null argument", index);
    const supra_bf_ctx* ctx = it.ctx;
    if (ctx->is3d) return fail(SUPRA_ERR_UNSUPPORTED, "batch item %u: 2D geometries only", index);
    if (it.rf_dtype != SUPRA_INT16 && it.rf_dtype != SUPRA_FLOAT32)
        return fail(SUPRA_ERR_UNSUPPORTED, "batch item %u: RF must be SUPRA_INT16 or SUPRA_FLOAT32", index);
    if (it.mode == SUPRA_BATCH_DAS) {
        if (it.out_dtype != SUPRA_INT16 && it.out_dtype != SUPRA_FLOAT32)
            return fail(SUPRA_ERR_UNSUPPORTED, "batch item %u: image must be SUPRA_INT16 or SUPRA_FLOAT32", index);
        if (it.window_type < SUPRA_WINDOW_RECTANGULAR || it.window_type > SUPRA_WINDOW_GAUSS)
            return fail(SUPRA_ERR_INVALID_ARG, "batch item %u: unknown window type", index);
        if (it.num_tx == 0 || it.num_channels == 0 || it.num_samples == 0)
            return fail(SUPRA_ERR_INVALID_ARG, "batch item %u: empty RF", index);
        if (static_cast<uint64_t>(it.num_channels) * it.num_samples * (it.rf_dtype == SUPRA_INT16 ? 2 : 4) >=
            (1ull << 31))
            return fail(SUPRA_ERR_TOO_LARGE, "batch item %u: one transmit's RF block must be below 2 GiB", index);
        if (!(it.sampling_frequency_hz > 0) || !(it.speed_of_sound_mm_per_s > 0) || !(it.f_number > 0) ||
            !std::isfinite(it.sampling_frequency_hz) || !std::isfinite(it.speed_of_sound_mm_per_s) ||
            !std::isfinite(it.f_number))
            return fail(SUPRA_ERR_INVALID_ARG, "batch item %u: fs, c and F# must be positive", index);
        if (!ctx->e2c.empty() && ctx->max_channel_in_map >= it.num_channels)
            return fail(SUPRA_ERR_INVALID_ARG, "batch item %u: element map exceeds the channels", index);
        for (uint32_t i = 0; i < ctx->n_rx; i++) {
            int32_t sel, mask;
            float w[4];
            scanline_selection(ctx->scan_mm[i], &sel, &mask, w);
            for (int k = 0; k < 4; k++) {
                const bool used = it.interpolate_transmits ? ((mask >> k) & 1) : (k == sel);
                if (used && ctx->scan_mm[i].tx[k].tx_scanline_idx >= it.num_tx)
                    return fail(SUPRA_ERR_INVALID_ARG, "batch item %u: scanline %u uses transmit %u of %u", index, i,
                                static_cast<unsigned>(ctx->scan_mm[i].tx[k].tx_scanline_idx), it.num_tx);
            }
        }
        return SUPRA_OK;
    }
    if (it.mode == SUPRA_BATCH_MV) {
        if (it.out_dtype != SUPRA_FLOAT32)
            return fail(SUPRA_ERR_UNSUPPORTED, "batch item %u: minimum variance writes SUPRA_FLOAT32", index);
        char what[32];
        std::snprintf(what, sizeof(what), "batch item %u", index);
        return check_rawdelay_call(ctx, what, it.rf_dtype, it.num_tx, it.num_channels, it.num_samples,
                                   it.sampling_frequency_hz, it.f_number, it.speed_of_sound_mm_per_s,
                                   it.window_type);
    }
    return fail(SUPRA_ERR_INVALID_ARG, "batch item %u: unknown mode %d", index, it.mode);
}

// DasArgs of frame f with its tables in the batch's table block.
DasArgs frame_das_args(const supra_batch* b, const FrameDesc& fd) {
    const supra_batch_item& it = fd.item;
    const supra_bf_ctx* ctx = it.ctx;
    char* base = static_cast<char*>(b->tables) + fd.table_offset;
    DasArgs a = rawdelay_args(ctx, it.rf, it.num_tx, it.num_channels, it.num_samples, fd.dt, it.f_number,
                              it.speed_of_sound_mm_per_s, it.window_type, it.window_parameter);
    a.scans = reinterpret_cast<const ScanlineDT*>(base);
    a.depths = reinterpret_cast<const float*>(base + align256(ctx->n_rx * sizeof(ScanlineDT)));
    a.ex = reinterpret_cast<const float*>(base + align256(ctx->n_rx * sizeof(ScanlineDT)) +
                                          align256(ctx->n_depths * sizeof(float)));
    a.ey = a.ex;    // 2D: unused
    a.e2c = ctx->d_e2c;
    a.out = it.out;
    a.offset = static_cast<float>(it.additional_offset);
    a.far_ok = ctx->far_ok_mm ? 1 : 0;
    return a;
}

template <typename T>
int upload(T** dst, const std::vector<T>& src, const char* what) {
    (void)hipFree(*dst);
    *dst = nullptr;
    if (src.empty()) return SUPRA_OK;
    hipError_t e = hipMalloc(dst, src.size() * sizeof(T));
    if (e == hipSuccess) e = hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, what);
    return SUPRA_OK;
}

int build_das_queue(supra_batch* b) {
    std::vector<DasBatchFrame> frames;
    std::vector<uint32_t> items;
    for (uint32_t f = 0; f < b->n_frames; f++) {
        const FrameDesc& fd = b->frames[f];
        if (fd.queue != 0) continue;
        DasBatchFrame fr;
        std::memset(&fr, 0, sizeof(fr));
        fr.a = frame_das_args(b, fd);
        fr.rf_dtype = fd.item.rf_dtype;
        fr.out_dtype = fd.item.out_dtype;
        fr.interp = fd.item.interpolate_transmits ? 1 : 0;
        fr.map = fd.item.ctx->d_e2c ? 1 : 0;
        fr.n_items = fr.a.n_rx * fr.a.depth_tiles;
        fr.frame = f;
        const uint32_t qf = static_cast<uint32_t>(frames.size());
        for (uint32_t i = 0; i < fr.n_items; i++) items.push_back(qf << 24 | i);
        frames.push_back(fr);
    }
    b->das_n_frames = static_cast<uint32_t>(frames.size());
    b->das_n_items = static_cast<uint32_t>(items.size());
    int st = upload(&b->das_frames, frames, "batch: DAS frames");
    if (st == SUPRA_OK) st = upload(&b->das_items, items, "batch: DAS items");
    return st;
}

// Plans the MV frames of queue q (1: L <= 32, 2: L <= 64) for an LDS budget.
int build_mv_queue(supra_batch* b, uint32_t q, size_t budget) {
    std::vector<MvBatchFrame> frames;
    std::vector<uint32_t> items;
    size_t lds = 0;
    for (uint32_t f = 0; f < b->n_frames; f++) {
        const FrameDesc& fd = b->frames[f];
        if (fd.queue != q) continue;
        const supra_batch_item& it = fd.item;
        MvBatchFrame fr;
        std::memset(&fr, 0, sizeof(fr));
        fr.a.g = frame_das_args(b, fd);
        fr.a.out = it.out;
        size_t frame_lds = 0;
        int st = mv_prepare(fr.a, it.ctx->n_rx, it.num_channels, it.ctx->n_depths, it.sub_array_size,
                            it.temporal_smoothing, it.sub_array_scaling_power, it.compute_means, it.loading_factor,
                            budget, &frame_lds);
        if (st != SUPRA_OK) return st;
        lds = std::max(lds, frame_lds);
        fr.rf_dtype = it.rf_dtype;
        fr.n_items = fr.a.n_rx * fr.a.n_tiles;
        if (fr.n_items >= kMaxItemsPerFrame) return fail(SUPRA_ERR_TOO_LARGE, "batch: frame %u too large", f);
        fr.frame = f;
        const uint32_t qf = static_cast<uint32_t>(frames.size());
        for (uint32_t i = 0; i < fr.n_items; i++) items.push_back(qf << 24 | i);
        frames.push_back(fr);
    }
    b->mv_n_frames[q - 1] = static_cast<uint32_t>(frames.size());
    b->mv_n_items[q - 1] = static_cast<uint32_t>(items.size());
    b->mv_lds[q - 1] = lds;
    int st = upload(&b->mv_frames[q - 1], frames, "batch: MV frames");
    if (st == SUPRA_OK) st = upload(&b->mv_items[q - 1], items, "batch: MV items");
    return st;
}

// LDS budgets of the MV queues: the whole LDS of a CU (one workgroup) or half
// of it (two), less room for the kernel's static LDS.
size_t mv_full_budget(const supra_batch* b) {
    int dev_lds = 65536;
    (void)hipDeviceGetAttribute(&dev_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, b->device);
    return static_cast<size_t>(dev_lds) - 256;
}
size_t mv_half_budget(const supra_batch* b) { return (mv_full_budget(b) + 256) / 2 - 256; }

BatchSync make_sync(const supra_batch* b, uint32_t q) {
    BatchSync s;
    s.head = b->heads + q;
    s.done = b->done;
    s.stamp = b->stamp;
    return s;
}

hipError_t launch_reset(const supra_batch* b, hipStream_t s) {
    const uint32_t n = std::max(b->n_frames + 1, kQueues);
    hipLaunchKernelGGL(batch_reset_kernel, dim3((n + 255) / 256), dim3(256), 0, s, b->heads, b->done, b->stamp,
                       b->n_frames);
    return hipGetLastError();
}

hipError_t launch_queue(const supra_batch* b, uint32_t q, hipStream_t s) {
    if (q == 0) {
        if (b->das_n_items == 0) return hipSuccess;
        DasBatchArgs a;
        a.frames = b->das_frames;
        a.items = b->das_items;
        a.n_items = b->das_n_items;
        a.sync = make_sync(b, 0);
        const uint32_t wg = std::min<uint32_t>(b->das_n_items, b->das_wg_per_cu * static_cast<uint32_t>(b->cus));
        return launch_das_batch(a, wg, s);
    }
    const uint32_t m = q - 1;
    if (b->mv_n_items[m] == 0) return hipSuccess;
    MvBatchArgs a;
    a.frames = b->mv_frames[m];
    a.items = b->mv_items[m];
    a.n_items = b->mv_n_items[m];
    a.sync = make_sync(b, q);
    const uint32_t wg = std::min<uint32_t>(b->mv_n_items[m], b->mv_wg_per_cu[m] * static_cast<uint32_t>(b->cus));
    return launch_mv_batch(a, q == 1 ? 32 : 64, wg, b->mv_lds[m], s);
}

// Average time of one launch of queue q alone (after a reset), in ms.
int time_queue(supra_batch* b, uint32_t q, float* ms) {
    hipStream_t s = b->streams[0];
    hipEvent_t e0, e1;
    hipError_t e;
    if ((e = hipEventCreate(&e0)) != hipSuccess) return hip_fail(e, "batch: event");
    if ((e = hipEventCreate(&e1)) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_fail(e, "batch: event");
    }
    float best = 1e30f;
    for (int rep = 0; rep < 3 && e == hipSuccess; rep++) {
        if ((e = launch_reset(b, s)) != hipSuccess) break;
        if ((e = hipEventRecord(e0, s)) != hipSuccess) break;
        if ((e = launch_queue(b, q, s)) != hipSuccess) break;
        if ((e = hipEventRecord(e1, s)) != hipSuccess) break;
        if ((e = hipEventSynchronize(e1)) != hipSuccess) break;
        float t = 0;
        if ((e = hipEventElapsedTime(&t, e0, e1)) != hipSuccess) break;
        if (rep > 0) best = std::min(best, t);    // first run warms caches and code
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return hip_fail(e, "batch: autotune run");
    *ms = best;
    return SUPRA_OK;
}

int autotune(supra_batch* b) {
    char line[160];
    int st;
    if (b->das_n_items) {
        const uint32_t cands[] = {2, 4, 7};
        float best = 1e30f;
        uint32_t pick = 4;
        for (uint32_t c : cands) {
            b->das_wg_per_cu = c;
            float t;
            if ((st = time_queue(b, 0, &t)) != SUPRA_OK) return st;
            std::snprintf(line, sizeof(line), "das wg/cu=%u %.3f ms; ", c, t);
            b->config += line;
            if (t < best) {
                best = t;
                pick = c;
            }
        }
        b->das_wg_per_cu = pick;
    }
    for (uint32_t q = 1; q <= 2; q++) {
        if (!b->mv_n_items[q - 1]) continue;
        struct Cand { size_t budget; uint32_t wg; };
        const Cand cands[] = {{mv_full_budget(b), 1}, {mv_half_budget(b), 2}};
        float best = 1e30f;
        Cand pick = cands[0];
        for (const Cand& c : cands) {
            if ((st = build_mv_queue(b, q, c.budget)) != SUPRA_OK) continue;   // does not fit: skip
            b->mv_wg_per_cu[q - 1] = c.wg;
            float t;
            if ((st = time_queue(b, q, &t)) != SUPRA_OK) return st;
            std::snprintf(line, sizeof(line), "mv%s lds=%zu wg/cu=%u %.3f ms; ", q == 1 ? "32" : "64", c.budget,
                          c.wg, t);
            b->config += line;
            if (t < best) {
                best = t;
                pick = c;
            }
        }
        clear_error();
        b->mv_budget[q - 1] = pick.budget;
        b->mv_wg_per_cu[q - 1] = pick.wg;
        if ((st = build_mv_queue(b, q, pick.budget)) != SUPRA_OK) return st;
    }
    return SUPRA_OK;
}

int capture_graph(supra_batch* b) {
    hipError_t e;
    hipStream_t s0 = b->streams[0];
    if ((e = hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return hip_fail(e, "batch: begin capture");
    e = launch_reset(b, s0);
    if (e == hipSuccess) e = hipEventRecord(b->ev_fork, s0);
    for (uint32_t q = 1; q < kQueues && e == hipSuccess; q++) e = hipStreamWaitEvent(b->streams[q], b->ev_fork, 0);
    for (uint32_t q = 0; q < kQueues && e == hipSuccess; q++) e = launch_queue(b, q, b->streams[q]);
    for (uint32_t q = 1; q < kQueues && e == hipSuccess; q++) e = hipEventRecord(b->ev_join[q - 1], b->streams[q]);
    for (uint32_t q = 1; q < kQueues && e == hipSuccess; q++) e = hipStreamWaitEvent(s0, b->ev_join[q - 1], 0);
    hipGraph_t g = nullptr;
    const hipError_t e_end = hipStreamEndCapture(s0, &g);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(e, "batch: capture");
    }
    if (e_end != hipSuccess) return hip_fail(e_end, "batch: end capture");
    b->graph = g;
    if ((e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0)) != hipSuccess)
        return hip_fail(e, "batch: instantiate");
    return SUPRA_OK;
}

}  // namespace

extern "C" int supra_batch_create(const supra_batch_item* items, uint32_t num_items, int32_t autotune_tiles,
                                  supra_batch** out_batch) {
    clear_error();
    if (!items || !out_batch || num_items == 0) return fail(SUPRA_ERR_INVALID_ARG, "batch: null or empty argument");
    *out_batch = nullptr;
    for (uint32_t i = 0; i < num_items; i++) {
        int st = check_item(items[i], i);
        if (st != SUPRA_OK) return st;
    }
    // The autotuner and the setup copies run on the batch's own streams: let
    // the caller's pending writes of RF and outputs finish first.
    hipError_t e0 = hipDeviceSynchronize();
    if (e0 != hipSuccess) return hip_fail(e0, "batch: synchronize");
    supra_batch* b = new (std::nothrow) supra_batch();
    if (!b) return fail(SUPRA_ERR_INVALID_ARG, "batch: out of host memory");
    std::unique_ptr<supra_batch> guard(b);
    hipError_t e = hipGetDevice(&b->device);
    if (e != hipSuccess) return fail(SUPRA_ERR_NO_DEVICE, "batch: no HIP device: %s", hipGetErrorString(e));
    int khz = 0;
    (void)hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->device);
    (void)hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device);
    b->clock_khz = khz > 0 ? khz : 100000.0;
    if (b->cus <= 0) b->cus = 1;
    b->n_frames = num_items;
    uint32_t per_queue[kQueues] = {0, 0, 0};
    size_t offset = 0;
    for (uint32_t i = 0; i < num_items; i++) {
        FrameDesc fd;
        fd.item = items[i];
        if (items[i].ctx->device != b->device)
            return fail(SUPRA_ERR_INVALID_ARG, "batch item %u: context of another device", i);
        if (items[i].mode == SUPRA_BATCH_DAS) {
            fd.queue = 0;
        } else {
            const uint32_t L = items[i].sub_array_size ? items[i].sub_array_size : items[i].num_channels / 2;
            fd.queue = L <= 32 ? 1 : 2;
        }
        if (++per_queue[fd.queue] > kMaxFramesPerQueue)
            return fail(SUPRA_ERR_TOO_LARGE, "batch: at most %u frames per kernel class", kMaxFramesPerQueue);
        fd.dt = 1.0 / items[i].sampling_frequency_hz;
        fd.table_offset = offset;
        offset += table_bytes(items[i].ctx);
        b->frames.push_back(fd);
    }
    // dt-space tables of every frame (host build, one upload)
    std::vector<char> host(offset);
    for (const FrameDesc& fd : b->frames) {
        supra_bf_ctx* ctx = fd.item.ctx;
        std::lock_guard<std::mutex> lock(ctx->mu);
        build_dt_tables(ctx, fd.dt, fd.item.speed_of_sound_mm_per_s);
        char* base = host.data() + fd.table_offset;
        std::memcpy(base, ctx->h_scans.data(), ctx->n_rx * sizeof(ScanlineDT));
        base += align256(ctx->n_rx * sizeof(ScanlineDT));
        std::memcpy(base, ctx->h_depths.data(), ctx->n_depths * sizeof(float));
        base += align256(ctx->n_depths * sizeof(float));
        std::memcpy(base, ctx->h_ex.data(), ctx->n_el * sizeof(float));
    }
    b->tables_bytes = offset;
    if ((e = hipMalloc(&b->tables, offset)) != hipSuccess ||
        (e = hipMemcpy(b->tables, host.data(), offset, hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "batch: tables");
    const size_t sync_bytes = align256((kQueues + num_items) * sizeof(uint32_t)) + (num_items + 1) * sizeof(uint64_t);
    if ((e = hipMalloc(&b->sync, sync_bytes)) != hipSuccess) return hip_fail(e, "batch: counters");
    b->heads = b->sync;
    b->done = b->sync + kQueues;
    b->stamp = reinterpret_cast<uint64_t*>(reinterpret_cast<char*>(b->sync) +
                                           align256((kQueues + num_items) * sizeof(uint32_t)));
    for (auto& s : b->streams)
        if ((e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess) return hip_fail(e, "batch: stream");
    if ((e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->ev_join[0], hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->ev_join[1], hipEventDisableTiming)) != hipSuccess)
        return hip_fail(e, "batch: event");
    int st = build_das_queue(b);
    for (uint32_t q = 1; q <= 2 && st == SUPRA_OK; q++) {
        st = build_mv_queue(b, q, mv_full_budget(b));
        b->mv_budget[q - 1] = mv_full_budget(b);
        b->mv_wg_per_cu[q - 1] = 1;
    }
    if (st != SUPRA_OK) return st;
    b->config.clear();
    if (autotune_tiles && (st = autotune(b)) != SUPRA_OK) return st;
    char line[200];
    std::snprintf(line, sizeof(line),
                  "chosen: das frames=%u items=%u wg/cu=%u | mv32 frames=%u items=%u wg/cu=%u lds=%zu | "
                  "mv64 frames=%u items=%u wg/cu=%u lds=%zu",
                  b->das_n_frames, b->das_n_items, b->das_wg_per_cu, b->mv_n_frames[0], b->mv_n_items[0],
                  b->mv_wg_per_cu[0], b->mv_lds[0], b->mv_n_frames[1], b->mv_n_items[1], b->mv_wg_per_cu[1],
                  b->mv_lds[1]);
    b->config += line;
    if ((e = hipDeviceSynchronize()) != hipSuccess) return hip_fail(e, "batch: setup");
    if ((st = capture_graph(b)) != SUPRA_OK) return st;
    *out_batch = guard.release();
    return SUPRA_OK;
}

extern "C" int supra_batch_launch(supra_batch* b, void* stream) {
    clear_error();
    if (!b) return fail(SUPRA_ERR_INVALID_ARG, "batch: null argument");
    const hipError_t e = hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "batch: graph launch");
    b->last_stream = static_cast<hipStream_t>(stream);
    return SUPRA_OK;
}

extern "C" int supra_batch_latencies(supra_batch* b, double* frame_us, double* total_us) {
    clear_error();
    if (!b) return fail(SUPRA_ERR_INVALID_ARG, "batch: null argument");
    hipError_t e = hipStreamSynchronize(b->last_stream);
    if (e != hipSuccess) return hip_fail(e, "batch: synchronize");
    std::vector<uint64_t> st(b->n_frames + 1);
    if ((e = hipMemcpy(st.data(), b->stamp, st.size() * sizeof(uint64_t), hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(e, "batch: timestamps");
    const uint64_t t0 = st[b->n_frames];
    double total = 0;
    for (uint32_t f = 0; f < b->n_frames; f++) {
        const double us = st[f] >= t0 ? static_cast<double>(st[f] - t0) / b->clock_khz * 1e3 : -1.0;
        if (frame_us) frame_us[f] = us;
        total = std::max(total, us);
    }
    if (total_us) *total_us = total;
    return SUPRA_OK;
}

extern "C" int supra_batch_config(const supra_batch* b, char* buffer, size_t buffer_size) {
    clear_error();
    if (!b || !buffer || buffer_size == 0) return fail(SUPRA_ERR_INVALID_ARG, "batch: null argument");
    std::snprintf(buffer, buffer_size, "%s", b->config.c_str());
    return SUPRA_OK;
}

extern "C" void supra_batch_destroy(supra_batch* b) { delete b; }
