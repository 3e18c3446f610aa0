// Host-callable launcher of the fused pre-delay + minimum-variance kernel
// (mv_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "bf_internal.hpp"

namespace supra_amd {

// The pre-delay of `g` (2D DasArgs as for the pre-delay kernel: one frame,
// n_ch output channels) feeding minimum variance over g.n_rx scanlines x
// g.n_depths samples; `out` is [depth][scanline].  Returns a SUPRA_* status
// and sets the thread's error message on failure.
int launch_mv_fused(const DasArgs& g, int32_t rf_dtype, uint32_t sub_array_size, uint32_t temporal_smoothing,
                    double sub_array_scaling_power, int32_t compute_means, double loading_factor, void* out,
                    int32_t out_dtype, hipStream_t stream);

}  // namespace supra_amd
