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

// Launch the 2D pre-delay kernel (output [frame][scanline][channel][depth]).
hipError_t launch_rawdelay(const DasArgs& a, int rf_dtype, int out_dtype, hipStream_t stream);

}  // namespace supra_amd
