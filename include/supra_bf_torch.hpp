/*
 * supra_bf_torch.hpp — libtorch (ROCm) autograd wrapper of the
 * differentiable 2D delay-and-sum: a torch::autograd::Function over
 * supra_das_forward_c / supra_das_adjoint (supra_bf.h).
 *
 * The reference links libtorch behind its SUPRA_TORCH switch
 * (src/SupraLib/CMakeLists.txt:342-356) for inference only; nothing in it
 * differentiates the beamformer.  This header is the C++ counterpart of
 * supra_amd.differentiable.DifferentiableDAS for libtorch programs:
 *
 *   supra_bf_ctx* ctx;  supra_bf_create(&geometry, &ctx);        // 2D geometry
 *   supra_amd::DasOptions o{40e6, 1.0, SUPRA_WINDOW_HANN, 0.f, false, 0};
 *   torch::Tensor img = supra_amd::das_forward_c(rf, c, ctx, o, num_depths, num_rx);
 *   img.pow(2).sum().backward();   // rf.grad(), c.grad()
 *
 * rf: float32 device tensor [frame][tx][channel][sample] (or [tx][ch][s]);
 * c: speed of sound in m/s, one value per frame (any floating type); the
 * image is [frame][depth][scanline] ([depth][scanline] for 3-D rf).  The
 * launch runs on the current HIP stream.  Header-only; link libsupra_bf.so
 * and libtorch.
 */
#ifndef SUPRA_BF_TORCH_HPP
#define SUPRA_BF_TORCH_HPP

#include <ATen/ATen.h>
#include <c10/hip/HIPStream.h>
#include <torch/csrc/autograd/custom_function.h>

#include <stdexcept>
#include <string>

#include "supra_bf.h"

namespace supra_amd {

struct DasOptions {
    double sampling_frequency_hz;
    double f_number;
    int32_t window_type;
    float window_parameter;
    bool interpolate_transmits;
    int32_t additional_offset;
};

namespace detail {

inline void check(int status, const char* what) {
    if (status != SUPRA_OK) throw std::runtime_error(std::string(what) + ": " + supra_last_error());
}

inline void* stream_of(const at::Tensor& t) {
    return c10::hip::getCurrentHIPStream(t.device().index()).stream();
}

}  // namespace detail

class DasCFunction : public torch::autograd::Function<DasCFunction> {
public:
    // rf [frame][tx][ch][s] float32, c [frame] (m/s)
    static at::Tensor forward(torch::autograd::AutogradContext* actx, const at::Tensor& rf, const at::Tensor& c,
                              int64_t handle, double fs_hz, double f_number, int64_t window_type,
                              double window_parameter, bool interpolate_transmits, int64_t additional_offset,
                              int64_t num_depths, int64_t num_rx) {
        TORCH_CHECK(rf.is_cuda() && rf.scalar_type() == at::kFloat && rf.dim() == 4,
                    "das_c: rf must be a float32 device tensor [frame][tx][ch][s]");
        TORCH_CHECK(c.dim() == 1 && c.size(0) == rf.size(0), "das_c: one speed of sound per frame");
        const at::Tensor x = rf.contiguous();
        const at::Tensor c_mm = (c.detach().to(at::kDouble) * 1000.0).to(rf.device()).contiguous();
        const int64_t frames = x.size(0), slots = interpolate_transmits ? 4 : 1;
        at::Tensor out = at::empty({frames, num_depths, num_rx}, x.options());
        at::Tensor norm = at::empty({frames, slots, num_rx, num_depths}, x.options());
        detail::check(supra_das_forward_c(reinterpret_cast<supra_bf_ctx*>(handle), static_cast<uint32_t>(frames),
                                          x.data_ptr<float>(), static_cast<uint32_t>(x.size(1)),
                                          static_cast<uint32_t>(x.size(2)), static_cast<uint32_t>(x.size(3)), fs_hz,
                                          f_number, c_mm.data_ptr<double>(), static_cast<int32_t>(window_type),
                                          static_cast<float>(window_parameter), interpolate_transmits ? 1 : 0,
                                          static_cast<int32_t>(additional_offset), out.data_ptr<float>(),
                                          norm.data_ptr<float>(), detail::stream_of(x)),
                      "supra_das_forward_c");
        actx->save_for_backward({x, c, norm});
        actx->saved_data["handle"] = handle;
        actx->saved_data["fs"] = fs_hz;
        actx->saved_data["fnum"] = f_number;
        actx->saved_data["window"] = window_type;
        actx->saved_data["wparam"] = window_parameter;
        actx->saved_data["interp"] = interpolate_transmits;
        actx->saved_data["offset"] = additional_offset;
        return out;
    }

    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* actx,
                                                   torch::autograd::variable_list grads) {
        const auto saved = actx->get_saved_variables();
        const at::Tensor& x = saved[0];
        const at::Tensor& c = saved[1];
        const at::Tensor& norm = saved[2];
        const at::Tensor gi = grads[0].to(at::kFloat).contiguous();
        const bool want_rf = actx->needs_input_grad(0), want_c = actx->needs_input_grad(1);
        const at::Tensor c_mm = (c.detach().to(at::kDouble) * 1000.0).to(x.device()).contiguous();
        at::Tensor grad_rf = want_rf ? at::empty_like(x) : at::Tensor();
        at::Tensor grad_c = want_c ? at::empty({x.size(0)}, x.options().dtype(at::kDouble)) : at::Tensor();
        detail::check(supra_das_adjoint(reinterpret_cast<supra_bf_ctx*>(actx->saved_data["handle"].toInt()),
                                        static_cast<uint32_t>(x.size(0)), x.data_ptr<float>(),
                                        static_cast<uint32_t>(x.size(1)), static_cast<uint32_t>(x.size(2)),
                                        static_cast<uint32_t>(x.size(3)), actx->saved_data["fs"].toDouble(),
                                        actx->saved_data["fnum"].toDouble(), c_mm.data_ptr<double>(),
                                        static_cast<int32_t>(actx->saved_data["window"].toInt()),
                                        static_cast<float>(actx->saved_data["wparam"].toDouble()),
                                        actx->saved_data["interp"].toBool() ? 1 : 0,
                                        static_cast<int32_t>(actx->saved_data["offset"].toInt()), norm.data_ptr<float>(),
                                        gi.data_ptr<float>(), want_rf ? grad_rf.data_ptr<float>() : nullptr,
                                        want_c ? grad_c.data_ptr<double>() : nullptr, detail::stream_of(x)),
                      "supra_das_adjoint");
        if (want_c) grad_c = (grad_c * 1000.0).to(c.scalar_type());   // 1/(mm/s) -> 1/(m/s)
        return {grad_rf, grad_c, at::Tensor(), at::Tensor(), at::Tensor(), at::Tensor(), at::Tensor(),
                at::Tensor(), at::Tensor(), at::Tensor(), at::Tensor()};
    }
};

// Image [frame][depth][scanline] (or [depth][scanline] for rf [tx][ch][s]).
inline at::Tensor das_forward_c(const at::Tensor& rf, const at::Tensor& c_m_per_s, supra_bf_ctx* ctx,
                                const DasOptions& o, int64_t num_depths, int64_t num_rx) {
    const bool single = rf.dim() == 3;
    const at::Tensor x = single ? rf.unsqueeze(0) : rf;
    at::Tensor c = c_m_per_s.dim() == 0 ? c_m_per_s.expand({x.size(0)}) : c_m_per_s;
    at::Tensor img = DasCFunction::apply(x, c, reinterpret_cast<int64_t>(ctx), o.sampling_frequency_hz, o.f_number,
                                         o.window_type, o.window_parameter, o.interpolate_transmits,
                                         o.additional_offset, num_depths, num_rx);
    return single ? img.squeeze(0) : img;
}

}  // namespace supra_amd

#endif /* SUPRA_BF_TORCH_HPP */
