// Host-callable launcher of the fused pre-delay + minimum-variance kernel
// (mv_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "bf_internal.hpp"

namespace supra_amd {

// Everything one minimum-variance launch needs (per frame in a batch).
struct MvArgs {
    const void* x;
    void* out;
    uint32_t n_rx, n_ch, n_d, L, T, K;
    uint32_t tile;          // output depths per workgroup (Dt)
    uint32_t n_tiles;
    uint32_t rs;            // LDS row stride of x (channels + padding)
    uint32_t packed;        // L (L + 1) / 2
    uint32_t waves;         // waves that solve (phase 3)
    uint32_t off_meta, off_mean, off_r, off_wave, wave_bytes;   // LDS byte offsets
    uint32_t off_fused;     // LDS byte offset of the pre-delay tables (fused mode)
    float p, alpha;
    int means;
    uint32_t pcg_iters;     // 0: Cholesky solve; else at most this many PCG iterations
    float pcg_threshold;    // PCG stops once |1 - R~ w| <= this
    DasArgs g;              // fused mode: geometry and RF of the pre-delay
};

struct MvBatchFrame {
    MvArgs a;                   // fused mode: a.g holds the pre-delay of the frame
    int32_t rf_dtype;           // (the queue's; kept for inspection)
    uint32_t n_items;           // work items (scanline x depth tile) of the frame
    uint32_t frame;             // index in the batch
    uint32_t pad;
};

struct MvBatchArgs {
    const MvBatchFrame* frames;
    const uint32_t* items;      // [item] = queue frame << 24 | item of the frame
    uint32_t n_items;
    BatchSync sync;
};

// sizeof(MvBatchFrame) as the kernels' translation unit sees it: a batch
// checks its own view against it before it uploads a frame table (objects
// built from different revisions of the struct must not be linked together).
size_t mv_batch_frame_bytes();

// Plan (tile of output depths, waves, LDS offsets) for a frame; lds_budget 0
// picks the single-launch default.  Returns a SUPRA_* status.
int mv_prepare(MvArgs& a, uint32_t num_scanlines, uint32_t num_channels, uint32_t num_samples,
               uint32_t sub_array_size, uint32_t temporal_smoothing, double sub_array_scaling_power,
               int32_t compute_means, double loading_factor, size_t lds_budget, size_t* lds_out);

// Persistent work-queue launch over a batch's minimum-variance frames of one
// kernel class (sub-array bound, RF type, mean removal); float output.
hipError_t launch_mv_batch(const MvBatchArgs& b, uint32_t lmax, int32_t rf_dtype, int32_t means, uint32_t workgroups,
                           size_t lds, hipStream_t stream);

// The pre-delay of `g` (2D DasArgs as for the pre-delay kernel: one frame,
// n_ch output channels) feeding minimum variance over g.n_rx scanlines x
// g.n_depths samples; `out` is [depth][scanline].  Returns a SUPRA_* status
// and sets the thread's error message on failure.
int launch_mv_fused(const DasArgs& g, int32_t rf_dtype, uint32_t sub_array_size, uint32_t temporal_smoothing,
                    double sub_array_scaling_power, int32_t compute_means, double loading_factor, void* out,
                    int32_t out_dtype, hipStream_t stream);

}  // namespace supra_amd
