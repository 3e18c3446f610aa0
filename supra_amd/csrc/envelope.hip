// Envelope detection and log compression of beamformed images, the stages
// after the beamformer in SUPRA's B-mode chain:
//   HilbertFirEnvelope (src/SupraLib/Beamformer/HilbertFirEnvelope.cu:28-104,
//     the FIR Hilbert transformer of FirFilterFactory.h:55-200, filter length
//     65 by default, HilbertFirEnvelopeNode.cpp:42)
//   LogCompressor (src/SupraLib/Beamformer/LogCompressor.cu:20-64,
//     LogCompressorNode.cpp:42-45).
//
// Envelope layout: one workgroup of 4 waves per (64 scanlines, 32 output
// samples, frame) of images [frame][sample][scanline]; lane = scanline, so
// every row of the window is one coalesced 256-byte load.  The workgroup
// stages the column window (32 + N - 1 rows) in LDS, loads issued eight
// rows at a time per wave before they are written, and each wave then
// computes every fourth output row from LDS.  Samples outside the image are
// staged as 0, which adds +-0 to the FIR sum where the reference skips them
// (the same value).  The FIR sum runs over the taps in ascending order in
// single precision without contraction, as the host restatement does, so
// the envelope is bit for bit the oracle's.  The fused variant applies the
// log compression to the float envelope in registers (the same operations
// as the two-step path with a float intermediate).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "status.hpp"
#include "supra_bf.h"

namespace supra_amd {
namespace {

constexpr int kEnvRows = 32;          // output samples per workgroup
constexpr int kEnvWaves = 4;          // waves per workgroup
constexpr int kMaxTaps = 257;         // HilbertFirEnvelopeNode's filterLength range (1..257)

struct EnvArgs {
    const void* in;
    void* out;
    uint32_t n_rx, n_s, n_taps, half;  // half = n_taps / 2 (window t - half .. t + half)
    uint32_t used_taps;               // taps that exist: min(n_taps, 2 half + 1)
    float in_scale, scale_den;        // log compression (fused variant)
    float h[kMaxTaps];
};

struct LogArgs {
    const void* in;
    void* out;
    uint64_t count;
    float in_scale, scale_den;
};

template <typename T> __device__ inline float load_as_float(const void* p, uint64_t i) {
    return static_cast<float>(static_cast<const T*>(p)[i]);
}

// clampCast (src/SupraLib/utilities/cudaUtility.h:111-147): fmaxf/fminf keep
// the non-NaN operand (NaN lands on the lower limit), then truncation.
template <typename T> __device__ inline T clamp_to(float x);
template <> __device__ inline float clamp_to<float>(float x) { return fminf(fmaxf(x, -FLT_MAX), FLT_MAX); }
template <> __device__ inline int16_t clamp_to<int16_t>(float x) {
    return static_cast<int16_t>(fminf(fmaxf(x, -32767.0f), 32767.0f));
}
template <> __device__ inline uint8_t clamp_to<uint8_t>(float x) {
    return static_cast<uint8_t>(fminf(fmaxf(x, 0.0f), 255.0f));
}

// thrustLogcompress::operator() (LogCompressor.cu:34-38).
__device__ inline float log_compress(float a, float in_scale, float scale_den) {
    return log10f(fabsf(a) * in_scale + 1.0f) * scale_den;
}

template <typename InT, typename OutT, bool LOG>
__global__ __launch_bounds__(64 * kEnvWaves) void envelope_kernel(const EnvArgs a) {
    extern __shared__ float win[];                 // [32 + 2 half rows][64 lanes]
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t s = blockIdx.x * 64u + lane;
    const uint32_t t0 = blockIdx.y * kEnvRows;
    const uint64_t frame = static_cast<uint64_t>(blockIdx.z) * a.n_s * a.n_rx;
    const uint32_t rows = kEnvRows + 2u * a.half;
    const bool col = s < a.n_rx;
    const InT* in = static_cast<const InT*>(a.in) + frame;
    for (uint32_t r0 = wave; r0 < rows; r0 += 8u * kEnvWaves) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const uint32_t r = r0 + u * kEnvWaves;
            const int64_t c = static_cast<int64_t>(t0) + r - a.half;
            v[u] = 0.0f;
            if (r < rows && col && c >= 0 && c < static_cast<int64_t>(a.n_s))
                v[u] = static_cast<float>(in[static_cast<uint64_t>(c) * a.n_rx + s]);
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const uint32_t r = r0 + u * kEnvWaves;
            if (r < rows) win[r * 64u + lane] = v[u];
        }
    }
    __syncthreads();
    if (!col) return;
    OutT* out = static_cast<OutT*>(a.out) + frame;
    for (uint32_t i = wave; i < kEnvRows; i += kEnvWaves) {
        const uint32_t t = t0 + i;
        if (t >= a.n_s) break;
        float acc = 0.0f;
#pragma unroll 8
        for (uint32_t j = 0; j < a.used_taps; j++) acc += win[(i + j) * 64u + lane] * a.h[j];
        const float v = win[(i + a.half) * 64u + lane];
        float env = sqrtf(v * v + acc * acc);
        // the fused path sees the envelope as the float image the two-node
        // chain passes on: clampCast<float> first (NaN -> -FLT_MAX, inf -> FLT_MAX)
        if (LOG) env = log_compress(clamp_to<float>(env), a.in_scale, a.scale_den);
        out[static_cast<uint64_t>(t) * a.n_rx + s] = clamp_to<OutT>(env);
    }
}

template <typename InT, typename OutT>
__global__ __launch_bounds__(256) void logcompress_kernel(const LogArgs a) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.count; i += stride)
        static_cast<OutT*>(a.out)[i] = clamp_to<OutT>(log_compress(load_as_float<InT>(a.in, i), a.in_scale, a.scale_den));
}

// FirFilterFactory::createFilter(N, FilterTypeHilbertTransformer,
// FilterWindowHamming): 2/(pi k) at odd offsets k from the centre, 0 at even
// ones, in double then float, times the float Hamming window.
void hilbert_taps(uint32_t n, float* h) {
    const int half = (static_cast<int>(n) - 1) / 2;
    const size_t max_n = n - 1;
    for (uint32_t i = 0; i < n; i++) {
        const int k = static_cast<int>(i) - half;
        h[i] = (k % 2 == 0) ? 0.0f : static_cast<float>(2.0 / (M_PI * k));
        h[i] *= static_cast<float>(0.54 - 0.46 * std::cos(2 * M_PI * i / max_n));
    }
}

// thrustLogcompress's constructor (LogCompressor.cu:28-31, :56-60): inMax is
// converted to the input element type first, outMax is 255 for both the
// float and the uint8 image.
void log_constants(int in_dtype, double dynamic_range_db, double gain, double in_max, float* in_scale,
                   float* scale_den) {
    const double dr = std::pow(10.0, dynamic_range_db / 20);
    double m = in_max;
    if (in_dtype == SUPRA_INT16) m = static_cast<int16_t>(in_max);
    else if (in_dtype == SUPRA_UINT8) m = static_cast<uint8_t>(in_max);
    else m = static_cast<float>(in_max);
    *in_scale = static_cast<float>(dr / m);
    *scale_den = static_cast<float>(gain * 255.0 / std::log10(dr + 1));
}

int check_types(const char* what, int in_dtype, int out_dtype, bool allow_u8_in, bool allow_i16_out, bool allow_u8_out) {
    const bool in_ok = in_dtype == SUPRA_FLOAT32 || in_dtype == SUPRA_INT16 || (allow_u8_in && in_dtype == SUPRA_UINT8);
    const bool out_ok = out_dtype == SUPRA_FLOAT32 || (allow_i16_out && out_dtype == SUPRA_INT16) ||
                        (allow_u8_out && out_dtype == SUPRA_UINT8);
    if (!in_ok || !out_ok) return fail(SUPRA_ERR_UNSUPPORTED, "%s: element types %d -> %d are not provided", what,
                                       in_dtype, out_dtype);
    return SUPRA_OK;
}

template <bool LOG, typename InT, typename OutT>
void launch_env_typed(const EnvArgs& a, dim3 grid, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((envelope_kernel<InT, OutT, LOG>), grid, dim3(64 * kEnvWaves), lds, s, a);
}

template <bool LOG>
int launch_envelope(const char* what, const void* image, int in_dtype, uint32_t n_rx, uint32_t n_s,
                    uint32_t n_frames, uint32_t filter_length, float in_scale, float scale_den, void* out,
                    int out_dtype, void* stream) {
    if (!image || !out) return fail(SUPRA_ERR_INVALID_ARG, "%s: null argument", what);
    if (n_frames > 65535u) return fail(SUPRA_ERR_TOO_LARGE, "%s: %u frames in one call (at most 65535)", what, n_frames);
    if (n_rx == 0 || n_s == 0 || n_frames == 0) return SUPRA_OK;
    if (filter_length < 1 || filter_length > kMaxTaps)
        return fail(SUPRA_ERR_INVALID_ARG, "%s: filter length %u outside [1, %d]", what, filter_length, kMaxTaps);
    int st = check_types(what, in_dtype, out_dtype, false, !LOG, LOG);
    if (st != SUPRA_OK) return st;
    if (static_cast<uint64_t>(n_s) * n_rx >= (1ull << 40)) return fail(SUPRA_ERR_TOO_LARGE, "%s: image too large", what);
    EnvArgs a;
    std::memset(&a, 0, sizeof(a));
    a.in = image;
    a.out = out;
    a.n_rx = n_rx;
    a.n_s = n_s;
    a.n_taps = filter_length;
    a.half = filter_length / 2;
    a.used_taps = std::min(filter_length, 2 * a.half + 1);
    a.in_scale = in_scale;
    a.scale_den = scale_den;
    hilbert_taps(filter_length, a.h);
    const dim3 grid((n_rx + 63) / 64, (n_s + kEnvRows - 1) / kEnvRows, n_frames);
    if (grid.y > 65535u) return fail(SUPRA_ERR_TOO_LARGE, "%s: too many samples per scanline", what);
    const size_t lds = static_cast<size_t>(kEnvRows + 2 * a.half) * 64 * sizeof(float);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (in_dtype == SUPRA_INT16) {
        if (out_dtype == SUPRA_FLOAT32) launch_env_typed<LOG, int16_t, float>(a, grid, lds, s);
        else if (out_dtype == SUPRA_INT16) launch_env_typed<LOG, int16_t, int16_t>(a, grid, lds, s);
        else launch_env_typed<LOG, int16_t, uint8_t>(a, grid, lds, s);
    } else {
        if (out_dtype == SUPRA_FLOAT32) launch_env_typed<LOG, float, float>(a, grid, lds, s);
        else if (out_dtype == SUPRA_INT16) launch_env_typed<LOG, float, int16_t>(a, grid, lds, s);
        else launch_env_typed<LOG, float, uint8_t>(a, grid, lds, s);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SUPRA_ERR_HIP, "%s: launch: %s", what, hipGetErrorString(e));
    return SUPRA_OK;
}

template <typename InT, typename OutT>
void launch_log_typed(const LogArgs& a, hipStream_t s) {
    const uint64_t blocks = std::min<uint64_t>((a.count + 255) / 256, 256ull * 64ull);
    hipLaunchKernelGGL((logcompress_kernel<InT, OutT>), dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, s, a);
}

}  // namespace
}  // namespace supra_amd

using namespace supra_amd;

extern "C" int supra_envelope_run(const void* image, int32_t in_dtype, uint32_t num_scanlines, uint32_t num_samples,
                                  uint32_t num_frames, uint32_t filter_length, void* out, int32_t out_dtype,
                                  void* stream) {
    clear_error();
    return launch_envelope<false>("envelope", image, in_dtype, num_scanlines, num_samples, num_frames, filter_length,
                                  0.0f, 0.0f, out, out_dtype, stream);
}

extern "C" int supra_logcompress_run(const void* in, int32_t in_dtype, uint64_t count, double dynamic_range_db,
                                     double gain, double in_max, void* out, int32_t out_dtype, void* stream) {
    clear_error();
    if (!in || !out) return fail(SUPRA_ERR_INVALID_ARG, "logcompress: null argument");
    int st = check_types("logcompress", in_dtype, out_dtype, true, false, true);
    if (st != SUPRA_OK) return st;
    if (count == 0) return SUPRA_OK;
    LogArgs a;
    a.in = in;
    a.out = out;
    a.count = count;
    log_constants(in_dtype, dynamic_range_db, gain, in_max, &a.in_scale, &a.scale_den);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool f32 = out_dtype == SUPRA_FLOAT32;
    if (in_dtype == SUPRA_INT16) f32 ? launch_log_typed<int16_t, float>(a, s) : launch_log_typed<int16_t, uint8_t>(a, s);
    else if (in_dtype == SUPRA_UINT8) f32 ? launch_log_typed<uint8_t, float>(a, s) : launch_log_typed<uint8_t, uint8_t>(a, s);
    else f32 ? launch_log_typed<float, float>(a, s) : launch_log_typed<float, uint8_t>(a, s);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SUPRA_ERR_HIP, "logcompress: launch: %s", hipGetErrorString(e));
    return SUPRA_OK;
}

extern "C" int supra_envelope_logcompress_run(const void* image, int32_t in_dtype, uint32_t num_scanlines,
                                              uint32_t num_samples, uint32_t num_frames, uint32_t filter_length,
                                              double dynamic_range_db, double gain, double in_max, void* out,
                                              int32_t out_dtype, void* stream) {
    clear_error();
    float in_scale, scale_den;
    log_constants(SUPRA_FLOAT32, dynamic_range_db, gain, in_max, &in_scale, &scale_den);
    return launch_envelope<true>("envelope + logcompress", image, in_dtype, num_scanlines, num_samples, num_frames,
                                 filter_length, in_scale, scale_den, out, out_dtype, stream);
}
