// Registers the differentiable 2D delay-and-sum of include/supra_bf_torch.hpp
// as the TorchScript operator supra_bf::das_c (libsupra_torch.so), so that
// libtorch and Python programs reach the same autograd Function through the
// dispatcher: torch.ops.supra_bf.das_c(rf, c, handle, fs_hz, f_number,
// window_type, window_parameter, interpolate_transmits, additional_offset,
// num_depths, num_rx), handle being a supra_bf_ctx* as an integer.
#include <torch/library.h>

#include "supra_bf_torch.hpp"

namespace {

at::Tensor das_c(const at::Tensor& rf, const at::Tensor& c, int64_t handle, double fs_hz, double f_number,
                 int64_t window_type, double window_parameter, bool interpolate_transmits, int64_t additional_offset,
                 int64_t num_depths, int64_t num_rx) {
    TORCH_CHECK(handle != 0, "das_c: null beamformer context");
    return supra_amd::DasCFunction::apply(rf, c, handle, fs_hz, f_number, window_type, window_parameter,
                                          interpolate_transmits, additional_offset, num_depths, num_rx);
}

}  // namespace

TORCH_LIBRARY(supra_bf, m) {
    m.def("das_c(Tensor rf, Tensor c, int handle, float fs_hz, float f_number, int window_type, "
          "float window_parameter, bool interpolate_transmits, int additional_offset, int num_depths, "
          "int num_rx) -> Tensor");
}

TORCH_LIBRARY_IMPL(supra_bf, CompositeImplicitAutograd, m) { m.impl("das_c", das_c); }
