// Host-callable launchers of the DAS kernels (das_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "bf_internal.hpp"

namespace supra_amd {

// Depth tile (lanes per workgroup) of the DAS launches.
uint32_t das_depth_tile();

// Launch the DAS kernel for rf/out element types SUPRA_INT16 / SUPRA_FLOAT32.
hipError_t launch_das(const DasArgs& a, int rf_dtype, int out_dtype, bool is3d, bool testsig, bool interp_tx,
                      bool map, hipStream_t stream);

// Launch the DelayAndStdDev / CoherenceFactorDelayAndSum kernels (kind
// SUPRA_BF_DELAY_AND_STD_DEV or SUPRA_BF_COHERENCE_FACTOR, das_variants.hip).
hipError_t launch_das_variant(const DasArgs& a, int rf_dtype, int out_dtype, bool is3d, int kind, bool interp_tx,
                              hipStream_t stream);

// Launch the 2D or 3D pre-delay kernel (output [frame][scanline][channel][depth]).
hipError_t launch_rawdelay(const DasArgs& a, int rf_dtype, int out_dtype, bool is3d, hipStream_t stream);

// Persistent work-queue DAS over the frames of a batch (das_kernels.hip).
hipError_t launch_das_batch(const DasBatchArgs& b, uint32_t workgroups, hipStream_t stream);

// sizeof(DasBatchFrame) as the kernels' translation unit sees it: a batch
// checks its own view against it before it uploads a frame table (objects
// built from different revisions of the struct must not be linked together).
size_t das_batch_frame_bytes();

// Differentiable 2D DAS (das_adjoint.hip).
hipError_t launch_frame_tables(const FrameTablesArgs& a, hipStream_t stream);
hipError_t launch_depth_table(const float* depths, const float* frame_c, float dt, float f_number, uint32_t n_d,
                              uint32_t frames, float* dtab, hipStream_t stream);
hipError_t launch_beta(const float* grad_image, const float* norm, float* beta, uint32_t n_rx, uint32_t n_d,
                       uint32_t slots, uint32_t frames, hipStream_t stream);
size_t das_adjoint_lds_bytes(uint32_t waves);
hipError_t launch_das_adjoint(const AdjointArgs& a, uint32_t frames, hipStream_t stream);
hipError_t launch_dc_reduce(const double* part, uint32_t per_frame, const double* c, double* grad_c, uint32_t frames,
                            hipStream_t stream);
hipError_t launch_zero_lines(float* grad_rf, const uint32_t* lines, uint32_t n_lines, uint32_t n_s,
                             uint64_t frame_stride, uint32_t frames, hipStream_t stream);
hipError_t launch_fill_zero(float* p, uint64_t n, hipStream_t stream);

}  // namespace supra_amd
