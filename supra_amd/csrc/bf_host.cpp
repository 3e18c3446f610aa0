// Host-side helpers of the engine: the apodization window table and the
// per-scanline transmit selection.
#include <algorithm>
#include <cmath>
#include <limits>

#include "bf_internal.hpp"
#include "status.hpp"

namespace supra_amd {

// WindowFunction::windowFunction<float> (src/SupraLib/Beamformer/WindowFunction.h:121-138).
// Evaluated in float with the same literals; Gauss goes through double as there.
static float window_value(int type, float r, float p) {
    const float two_pi = 2 * static_cast<float>(M_PI);
    switch (type) {
    case SUPRA_WINDOW_RECTANGULAR:
        return 1.0f;
    case SUPRA_WINDOW_HANN:
        return (1 - p) * (0.5f - 0.5f * std::cos(two_pi * ((r + 1) * 0.5f))) + p;
    case SUPRA_WINDOW_HAMMING:
        return (1 - p) * (0.54f - 0.46f * std::cos(two_pi * ((r + 1) * 0.5f))) + p;
    case SUPRA_WINDOW_GAUSS: {
        const float q = r / p;
        return static_cast<float>(1.0 / (p * std::sqrt(2.0 * M_PI)) * std::exp((-1.0 / 2.0) * q * q));
    }
    default:
        return 0.0f;
    }
}

// WindowFunction::WindowFunction (WindowFunction.cpp:18-47): sample the window
// at r_i = (i / (N-1)) * 2 - 1 and normalise by the table maximum.
void build_window_lut(int window_type, float window_parameter, float* lut) {
    const int n = kWindowEntries;
    float vmax = std::numeric_limits<float>::min();
    for (int i = 0; i < n; i++) {
        float r = (static_cast<float>(i) / static_cast<float>(n - 1)) * 2 - 1;
        float v = window_value(window_type, r, window_parameter);
        lut[i] = v;
        vmax = std::max(vmax, v);
    }
    for (int i = 0; i < n; i++) lut[i] /= vmax;
}

// Transmit choice of the DAS/RawDelay kernels (RxBeamformerCuda.cu:274-295,
// RawDelay.cu:438-451): running maximum kept in float, compared against the
// double weights; ties keep the lowest index.  Returns false if a weighted
// transmit is out of range is left to the caller.
bool scanline_selection(const supra_scanline_rx& s, int32_t* sel, int32_t* active_mask, float txw[4]) {
    int best = 0;
    float best_w = static_cast<float>(s.tx_weights[0]);
    for (int k = 1; k < 4; k++) {
        if (s.tx_weights[k] > best_w) {
            best_w = static_cast<float>(s.tx_weights[k]);
            best = k;
        }
    }
    *sel = (s.tx_weights[best] > 0.0) ? best : -1;
    int mask = 0;
    for (int k = 0; k < 4; k++) {
        txw[k] = static_cast<float>(s.tx_weights[k]);
        if (s.tx_weights[k] > 0.0) mask |= 1 << k;
    }
    *active_mask = mask;
    return true;
}

}  // namespace supra_amd

extern "C" int supra_window_lut(int32_t window_type, float window_parameter, float* lut_out) {
    using namespace supra_amd;
    clear_error();
    if (!lut_out) return fail(SUPRA_ERR_INVALID_ARG, "window: null output");
    if (window_type < SUPRA_WINDOW_RECTANGULAR || window_type > SUPRA_WINDOW_GAUSS)
        return fail(SUPRA_ERR_INVALID_ARG, "window: unknown window type %d", window_type);
    build_window_lut(window_type, window_parameter, lut_out);
    return SUPRA_OK;
}
