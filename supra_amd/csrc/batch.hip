// Batched beamforming of many acquisitions of uneven size in one graph
// launch (no reference counterpart: SUPRA processes one USRawData per node
// call, BeamformingNode.cpp:128-177, and one RawDelay + MV chain per frame).
//
// A batch is planned once: every acquisition becomes a frame of work items
// (scanline x tile of output depths) in a queue of its kernel class:
// delay-and-sum (one queue, the RF/image types and options are per-frame
// branches), or minimum variance with the pre-delay fused in, one queue per
// (L <= 32 or L <= 64, RF type, mean removal) so that each runs a kernel
// specialised for it.  Each queue is served by a persistent work-queue
// kernel sized to the GPU; workgroups take items in frame order, so frames
// complete in submission order within a queue, and the last item of a frame
// stamps the frame's completion time from the device wall clock.  The queues
// run concurrently (a fork/join over one stream per queue) or one after
// another (see the autotuner below), and the whole
// launch — reset of the queue counters, the kernels — is captured once in a
// hipGraph; supra_batch_launch replays it.
//
// Footprints differ by an order of magnitude (DAS: 70 VGPRs, 10 KB LDS; MV:
// 256 VGPRs and up to the whole LDS), so one kernel for all items would run
// the delay-and-sum items at the minimum-variance occupancy; the queues keep
// each class at its own.
//
// The autotuner times each queue alone on the real batch for a few
// candidates (DAS: 2, 4, 7 or 8 persistent workgroups per CU; MV: one to
// four workgroups per CU, each with the depth tile that fits its share of
// the LDS) and keeps the fastest; then it
// times the whole launch with the queues concurrent and with the queues one
// after another (shortest time per frame first) and keeps the second unless
// the first is more than 5 % faster.  Without the autotuner the queues run
// concurrently.
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

constexpr uint32_t kMaxQueues = 9;         // DAS + MV x {L<=32, L<=64} x {int16, float} x {means}
constexpr uint32_t kMaxFramesPerQueue = 256;
constexpr uint32_t kMaxItemsPerFrame = 1u << 24;

__global__ void batch_reset_kernel(uint32_t* heads, uint32_t* done, uint64_t* stamp, uint32_t n_frames) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kMaxQueues) heads[i] = 0;
    if (i < n_frames) {
        done[i] = 0;
        stamp[i] = 0;
    }
    if (i == 0) stamp[n_frames] = wall_clock64();   // launch start
}

struct FrameDesc {
    supra_batch_item item;
    uint32_t queue;            // index into supra_batch::queues
    double dt;
    size_t table_offset;       // bytes into the table block
};

struct Queue {
    bool mv = false;
    uint32_t lmax = 0;         // MV: 32 or 64
    int32_t rf_dtype = 0;      // MV
    int32_t means = 0;         // MV
    void* frames = nullptr;    // DasBatchFrame[] or MvBatchFrame[] (device)
    uint32_t* items = nullptr;
    uint32_t n_items = 0, n_frames = 0;
    size_t lds = 0, budget = 0;
    uint32_t wg_per_cu = 1;
    float alone_ms = 0;        // autotuned time of the queue alone (0: not timed)
    hipStream_t stream = nullptr;
    hipEvent_t joined = nullptr;
};

}  // namespace

struct supra_batch {
    int device = 0;
    uint32_t n_frames = 0;
    int cus = 0;
    double clock_khz = 0;
    std::vector<FrameDesc> frames;
    std::vector<Queue> queues;
    // device memory
    void* tables = nullptr;            // per-frame dt-space tables
    uint32_t* sync = nullptr;          // heads[kMaxQueues] + done[n_frames] (uint32) + stamps (uint64)
    uint64_t* stamp = nullptr;
    bool launched = false;             // supra_batch_launch has run (the stamps are meaningful)
    uint32_t* heads = nullptr;
    uint32_t* done = nullptr;
    std::string config;
    // schedule: the queues concurrently (fork/join), or one after another
    // in `order` on one stream (chosen by the autotuner)
    bool ordered = false;
    std::vector<uint32_t> order;
    // graph
    hipStream_t origin = nullptr;      // capture stream: reset, fork, join
    hipEvent_t ev_fork = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t last_stream = nullptr;

    ~supra_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (origin) (void)hipStreamDestroy(origin);
        for (Queue& q : queues) {
            if (q.joined) (void)hipEventDestroy(q.joined);
            if (q.stream) (void)hipStreamDestroy(q.stream);
            (void)hipFree(q.frames);
            (void)hipFree(q.items);
        }
        (void)hipFree(tables);
        (void)hipFree(sync);
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
                                   it.window_type, false);
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

template <typename T, typename D>
int upload(D** dst, const std::vector<T>& src, const char* what) {
    (void)hipFree(*dst);
    *dst = nullptr;
    if (src.empty()) return SUPRA_OK;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, src.size() * sizeof(T));
    if (e == hipSuccess) e = hipMemcpy(p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(p);
        return hip_fail(e, what);
    }
    *dst = static_cast<D*>(p);
    return SUPRA_OK;
}

int build_das_queue(supra_batch* b, uint32_t qi) {
    if (das_batch_frame_bytes() != sizeof(DasBatchFrame))
        return fail(SUPRA_ERR_INVALID_ARG, "batch: DAS frame layout %zu bytes here, %zu in the kernels: the library "
                    "links objects of different revisions, rebuild it", sizeof(DasBatchFrame), das_batch_frame_bytes());
    Queue& q = b->queues[qi];
    std::vector<DasBatchFrame> frames;
    std::vector<uint32_t> items;
    for (uint32_t f = 0; f < b->n_frames; f++) {
        const FrameDesc& fd = b->frames[f];
        if (fd.queue != qi) continue;
        DasBatchFrame fr;
        std::memset(&fr, 0, sizeof(fr));
        fr.a = frame_das_args(b, fd);
        fr.rf_dtype = fd.item.rf_dtype;
        fr.out_dtype = fd.item.out_dtype;
        fr.interp = fd.item.interpolate_transmits ? 1 : 0;
        fr.map = fd.item.ctx->d_e2c ? 1 : 0;
        fr.n_items = fr.a.n_rx * fr.a.depth_tiles;
        if (fr.n_items >= kMaxItemsPerFrame) return fail(SUPRA_ERR_TOO_LARGE, "batch: frame %u too large", f);
        fr.frame = f;
        const uint32_t qf = static_cast<uint32_t>(frames.size());
        for (uint32_t i = 0; i < fr.n_items; i++) items.push_back(qf << 24 | i);
        frames.push_back(fr);
    }
    q.n_frames = static_cast<uint32_t>(frames.size());
    q.n_items = static_cast<uint32_t>(items.size());
    int st = upload(&q.frames, frames, "batch: DAS frames");
    if (st == SUPRA_OK) st = upload(&q.items, items, "batch: DAS items");
    return st;
}

// Plans the MV frames of queue qi for an LDS budget.
int build_mv_queue(supra_batch* b, uint32_t qi, size_t budget) {
    if (mv_batch_frame_bytes() != sizeof(MvBatchFrame))
        return fail(SUPRA_ERR_INVALID_ARG, "batch: MV frame layout %zu bytes here, %zu in the kernels: the library "
                    "links objects of different revisions, rebuild it", sizeof(MvBatchFrame), mv_batch_frame_bytes());
    Queue& q = b->queues[qi];
    std::vector<MvBatchFrame> frames;
    std::vector<uint32_t> items;
    size_t lds = 0;
    for (uint32_t f = 0; f < b->n_frames; f++) {
        const FrameDesc& fd = b->frames[f];
        if (fd.queue != qi) continue;
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
    q.n_frames = static_cast<uint32_t>(frames.size());
    q.n_items = static_cast<uint32_t>(items.size());
    q.lds = lds;
    q.budget = budget;
    int st = upload(&q.frames, frames, "batch: MV frames");
    if (st == SUPRA_OK) st = upload(&q.items, items, "batch: MV items");
    return st;
}

// LDS budgets of the MV queues: the whole LDS of a CU (one workgroup) or half
// of it (two), less room for the kernel's static LDS (16 bytes: the claimed
// item's broadcast slot).
constexpr size_t kMvStaticLdsReserve = 64;
size_t mv_full_budget(const supra_batch* b) {
    int dev_lds = 65536;
    (void)hipDeviceGetAttribute(&dev_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, b->device);
    return static_cast<size_t>(dev_lds) - kMvStaticLdsReserve;
}
// The LDS budget of one of `wg` workgroups sharing a CU's LDS.
size_t mv_share_budget(const supra_batch* b, uint32_t wg) {
    return (mv_full_budget(b) + kMvStaticLdsReserve) / wg - kMvStaticLdsReserve;
}

BatchSync make_sync(const supra_batch* b, uint32_t q) {
    BatchSync s;
    s.head = b->heads + q;
    s.done = b->done;
    s.stamp = b->stamp;
    return s;
}

hipError_t launch_reset(const supra_batch* b, hipStream_t s) {
    const uint32_t n = std::max(b->n_frames + 1, kMaxQueues);
    hipLaunchKernelGGL(batch_reset_kernel, dim3((n + 255) / 256), dim3(256), 0, s, b->heads, b->done, b->stamp,
                       b->n_frames);
    return hipGetLastError();
}

hipError_t launch_queue(const supra_batch* b, uint32_t qi, hipStream_t s) {
    const Queue& q = b->queues[qi];
    if (q.n_items == 0) return hipSuccess;
    const uint32_t wg = std::min<uint32_t>(q.n_items, q.wg_per_cu * static_cast<uint32_t>(b->cus));
    if (!q.mv) {
        DasBatchArgs a;
        a.frames = static_cast<const DasBatchFrame*>(q.frames);
        a.items = q.items;
        a.n_items = q.n_items;
        a.sync = make_sync(b, qi);
        return launch_das_batch(a, wg, s);
    }
    MvBatchArgs a;
    a.frames = static_cast<const MvBatchFrame*>(q.frames);
    a.items = q.items;
    a.n_items = q.n_items;
    a.sync = make_sync(b, qi);
    return launch_mv_batch(a, q.lmax, q.rf_dtype, q.means, wg, q.lds, s);
}

// Average time of one launch of queue q alone (after a reset), in ms.
int time_queue(supra_batch* b, uint32_t q, float* ms) {
    hipStream_t s = b->origin;
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

// One launch of every queue (after a reset) on the origin stream: all
// concurrently through the fork/join streams, or one after another in
// b->order.  Returns the time of the best of two runs after a warm-up.
int time_schedule(supra_batch* b, bool ordered, float* ms) {
    hipStream_t s0 = b->origin;
    hipEvent_t e0, e1;
    hipError_t e;
    if ((e = hipEventCreate(&e0)) != hipSuccess) return hip_fail(e, "batch: event");
    if ((e = hipEventCreate(&e1)) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_fail(e, "batch: event");
    }
    float best = 1e30f;
    for (int rep = 0; rep < 3 && e == hipSuccess; rep++) {
        if ((e = launch_reset(b, s0)) != hipSuccess) break;
        if ((e = hipEventRecord(e0, s0)) != hipSuccess) break;
        if (ordered) {
            for (uint32_t qi : b->order)
                if ((e = launch_queue(b, qi, s0)) != hipSuccess) break;
        } else {
            if ((e = hipEventRecord(b->ev_fork, s0)) != hipSuccess) break;
            for (uint32_t qi = 0; qi < b->queues.size() && e == hipSuccess; qi++) {
                e = hipStreamWaitEvent(b->queues[qi].stream, b->ev_fork, 0);
                if (e == hipSuccess) e = launch_queue(b, qi, b->queues[qi].stream);
                if (e == hipSuccess) e = hipEventRecord(b->queues[qi].joined, b->queues[qi].stream);
                if (e == hipSuccess) e = hipStreamWaitEvent(s0, b->queues[qi].joined, 0);
            }
        }
        if (e != hipSuccess) break;
        if ((e = hipEventRecord(e1, s0)) != hipSuccess) break;
        if ((e = hipEventSynchronize(e1)) != hipSuccess) break;
        float t = 0;
        if ((e = hipEventElapsedTime(&t, e0, e1)) != hipSuccess) break;
        if (rep > 0) best = std::min(best, t);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return hip_fail(e, "batch: schedule run");
    *ms = best;
    return SUPRA_OK;
}

// The order that minimises the mean frame latency when the queues run one
// after another (Smith's rule: ascending queue time per frame), and the
// schedule: one after another unless running the queues concurrently is
// more than 5 % faster.  One after another a short queue's frames finish as
// soon as it does instead of sharing the GPU with the long queues.
int choose_schedule(supra_batch* b) {
    b->order.clear();
    for (uint32_t qi = 0; qi < b->queues.size(); qi++)
        if (b->queues[qi].n_items) b->order.push_back(qi);
    std::stable_sort(b->order.begin(), b->order.end(), [&](uint32_t x, uint32_t y) {
        const Queue &qx = b->queues[x], &qy = b->queues[y];
        return qx.alone_ms * qy.n_frames < qy.alone_ms * qx.n_frames;
    });
    float t_conc = 0, t_ord = 0;
    int st = time_schedule(b, false, &t_conc);
    if (st == SUPRA_OK) st = time_schedule(b, true, &t_ord);
    if (st != SUPRA_OK) return st;
    b->ordered = t_ord <= 1.05f * t_conc;
    char line[160];
    std::snprintf(line, sizeof(line), "concurrent %.3f ms; in order %.3f ms; ", t_conc, t_ord);
    b->config += line;
    return SUPRA_OK;
}

int autotune(supra_batch* b) {
    char line[200];
    int st;
    for (uint32_t qi = 0; qi < b->queues.size(); qi++) {
        Queue& q = b->queues[qi];
        if (!q.n_items) continue;
        if (!q.mv) {
            const uint32_t cands[] = {2, 4, 7, 8};
            float best = 1e30f;
            uint32_t pick = 4;
            for (uint32_t c : cands) {
                q.wg_per_cu = c;
                float t;
                if ((st = time_queue(b, qi, &t)) != SUPRA_OK) return st;
                std::snprintf(line, sizeof(line), "das wg/cu=%u %.3f ms; ", c, t);
                b->config += line;
                if (t < best) {
                    best = t;
                    pick = c;
                }
            }
            q.wg_per_cu = pick;
            q.alone_ms = best;
            continue;
        }
        struct Cand { size_t budget; uint32_t wg; };
        // one to four workgroups per CU, each planning its depth tile for
        // its share of the LDS (a larger tile re-reads less of the temporal
        // halo; more workgroups hide more latency)
        const Cand cands[] = {{mv_full_budget(b), 1}, {mv_share_budget(b, 2), 2}, {mv_share_budget(b, 3), 3},
                              {mv_share_budget(b, 4), 4}};
        float best = 1e30f;
        Cand pick = cands[0];
        for (const Cand& c : cands) {
            if (build_mv_queue(b, qi, c.budget) != SUPRA_OK) continue;   // does not fit: skip
            q.wg_per_cu = c.wg;
            float t;
            if ((st = time_queue(b, qi, &t)) != SUPRA_OK) return st;
            std::snprintf(line, sizeof(line), "mv L<=%u %s%s lds=%zu wg/cu=%u %.3f ms; ", q.lmax,
                          q.rf_dtype == SUPRA_INT16 ? "int16" : "float", q.means ? " means" : "", c.budget, c.wg, t);
            b->config += line;
            if (t < best) {
                best = t;
                pick = c;
            }
        }
        clear_error();
        q.wg_per_cu = pick.wg;
        q.alone_ms = best;
        if ((st = build_mv_queue(b, qi, pick.budget)) != SUPRA_OK) return st;
    }
    return choose_schedule(b);
}

int capture_graph(supra_batch* b) {
    hipError_t e;
    hipStream_t s0 = b->origin;
    if ((e = hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return hip_fail(e, "batch: begin capture");
    e = launch_reset(b, s0);
    if (b->ordered) {
        for (uint32_t q : b->order)
            if (e == hipSuccess) e = launch_queue(b, q, s0);
    } else {
        if (e == hipSuccess) e = hipEventRecord(b->ev_fork, s0);
        for (uint32_t q = 0; q < b->queues.size() && e == hipSuccess; q++) {
            e = hipStreamWaitEvent(b->queues[q].stream, b->ev_fork, 0);
            if (e == hipSuccess) e = launch_queue(b, q, b->queues[q].stream);
            if (e == hipSuccess) e = hipEventRecord(b->queues[q].joined, b->queues[q].stream);
            if (e == hipSuccess) e = hipStreamWaitEvent(s0, b->queues[q].joined, 0);
        }
    }
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
    std::vector<uint32_t> per_queue;
    size_t offset = 0;
    for (uint32_t i = 0; i < num_items; i++) {
        FrameDesc fd;
        fd.item = items[i];
        if (items[i].ctx->device != b->device)
            return fail(SUPRA_ERR_INVALID_ARG, "batch item %u: context of another device", i);
        Queue key;
        if (items[i].mode == SUPRA_BATCH_MV) {
            const uint32_t L = items[i].sub_array_size ? items[i].sub_array_size : items[i].num_channels / 2;
            key.mv = true;
            key.lmax = L <= 32 ? 32 : 64;
            key.rf_dtype = items[i].rf_dtype;
            key.means = items[i].compute_means ? 1 : 0;
        }
        uint32_t qi = 0;
        while (qi < b->queues.size() && !(b->queues[qi].mv == key.mv && b->queues[qi].lmax == key.lmax &&
                                          b->queues[qi].rf_dtype == key.rf_dtype && b->queues[qi].means == key.means))
            qi++;
        if (qi == b->queues.size()) {
            if (qi == kMaxQueues) return fail(SUPRA_ERR_TOO_LARGE, "batch: too many kernel classes");
            b->queues.push_back(key);
            per_queue.push_back(0);
        }
        if (++per_queue[qi] > kMaxFramesPerQueue)
            return fail(SUPRA_ERR_TOO_LARGE, "batch: at most %u frames per kernel class", kMaxFramesPerQueue);
        fd.queue = qi;
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
    if ((e = hipMalloc(&b->tables, offset)) != hipSuccess ||
        (e = hipMemcpy(b->tables, host.data(), offset, hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "batch: tables");
    const size_t sync_bytes =
        align256((kMaxQueues + num_items) * sizeof(uint32_t)) + (num_items + 1) * sizeof(uint64_t);
    if ((e = hipMalloc(&b->sync, sync_bytes)) != hipSuccess) return hip_fail(e, "batch: counters");
    // (not captured: the graph clears the counters with its reset kernel)
    if ((e = hipMemset(b->sync, 0, sync_bytes)) != hipSuccess) return hip_fail(e, "batch: counters");
    b->heads = b->sync;
    b->done = b->sync + kMaxQueues;
    b->stamp = reinterpret_cast<uint64_t*>(reinterpret_cast<char*>(b->sync) +
                                           align256((kMaxQueues + num_items) * sizeof(uint32_t)));
    if ((e = hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming)) != hipSuccess)
        return hip_fail(e, "batch: stream");
    for (Queue& q : b->queues)
        if ((e = hipStreamCreateWithFlags(&q.stream, hipStreamNonBlocking)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&q.joined, hipEventDisableTiming)) != hipSuccess)
            return hip_fail(e, "batch: stream");
    int st = SUPRA_OK;
    for (uint32_t qi = 0; qi < b->queues.size() && st == SUPRA_OK; qi++) {
        Queue& q = b->queues[qi];
        if (!q.mv) {
            q.wg_per_cu = 4;
            st = build_das_queue(b, qi);
        } else {
            q.wg_per_cu = 1;
            st = build_mv_queue(b, qi, mv_full_budget(b));
        }
    }
    if (st != SUPRA_OK) return st;
    b->config.clear();
    if (autotune_tiles && (st = autotune(b)) != SUPRA_OK) return st;
    b->config += b->ordered ? "chosen: in order" : "chosen: concurrent";
    for (const Queue& q : b->queues) {
        char line[200];
        if (!q.mv)
            std::snprintf(line, sizeof(line), " [das frames=%u items=%u wg/cu=%u]", q.n_frames, q.n_items, q.wg_per_cu);
        else
            std::snprintf(line, sizeof(line), " [mv L<=%u %s%s frames=%u items=%u wg/cu=%u lds=%zu]", q.lmax,
                          q.rf_dtype == SUPRA_INT16 ? "int16" : "float", q.means ? " means" : "", q.n_frames,
                          q.n_items, q.wg_per_cu, q.lds);
        b->config += line;
    }
    if (b->ordered) {
        b->config += " order:";
        for (uint32_t q : b->order) b->config += " " + std::to_string(q);
    }
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
    b->launched = true;
    return SUPRA_OK;
}

extern "C" int supra_batch_latencies(supra_batch* b, double* frame_us, double* total_us) {
    clear_error();
    if (!b) return fail(SUPRA_ERR_INVALID_ARG, "batch: null argument");
    if (!b->launched) return fail(SUPRA_ERR_INVALID_ARG, "batch: latencies before the first supra_batch_launch");
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
