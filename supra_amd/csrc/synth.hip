// Seeded synthetic RF (point scatterers, Gaussian pulse, white noise) on the
// host and on the GPU.  Test and benchmark input only; see supra_synth.h.
//
// Echo sums are accumulated in 32-bit fixed point so that the result does not
// depend on the order in which threads add scatterers, and the signal power
// is reduced per line and summed on the host in line order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "status.hpp"
#include "supra_bf.h"
#include "supra_synth.h"

namespace supra_amd {
namespace {

constexpr double kFixedScale = 1048576.0;   // 2^20
constexpr uint32_t kMaxDeviceSamples = 8192;

__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__host__ __device__ inline double uniform01(uint64_t seed, uint64_t index) {
    uint64_t h = mix64(seed * 0x9e3779b97f4a7c15ull + mix64(index + 0x632be59bd9b4e019ull));
    return static_cast<double>(h >> 11) * (1.0 / 9007199254740992.0);
}

// Standard normal from two counter draws (Box-Muller), single precision.
__host__ __device__ inline float gauss(uint64_t seed, uint64_t index) {
    float u1 = static_cast<float>(1.0 - uniform01(seed, 2 * index));       // (0, 1]
    float u2 = static_cast<float>(uniform01(seed, 2 * index + 1));
    return sqrtf(-2.0f * logf(u1)) * cosf(6.2831853f * u2);
}

struct LineGeom {
    double ox, oy, oz, dx, dy, dz, t0;
    double ex, ey, ez;
    bool has_element;
};

__host__ __device__ inline LineGeom line_geometry(const supra_synth_params& p, const double* tx_origin,
                                                  const double* tx_dir, const double* tx_t0, const uint32_t* first,
                                                  const uint32_t* last, const double* elem, uint32_t tx, uint32_t ch) {
    LineGeom g;
    g.ox = tx_origin[3 * tx];
    g.oy = tx_origin[3 * tx + 1];
    g.oz = tx_origin[3 * tx + 2];
    g.dx = tx_dir[3 * tx];
    g.dy = tx_dir[3 * tx + 1];
    g.dz = tx_dir[3 * tx + 2];
    g.t0 = tx_t0[tx];
    uint32_t e;
    if (p.num_channels >= p.num_elements) {
        e = ch;
        g.has_element = ch < p.num_elements;
    } else {
        uint32_t f = first[tx];
        e = f + (ch + p.num_channels - f % p.num_channels) % p.num_channels;
        g.has_element = e <= last[tx] && e < p.num_elements;
    }
    if (g.has_element) {
        g.ex = elem[3 * e];
        g.ey = elem[3 * e + 1];
        g.ez = elem[3 * e + 2];
    } else {
        g.ex = g.ey = g.ez = 0;
    }
    return g;
}

struct PulseShape {
    double fs, c_mm, f0, sigma_t, half_window, beam2;
};

__host__ __device__ inline PulseShape pulse_shape(const supra_synth_params& p) {
    PulseShape s;
    s.fs = p.sampling_frequency_hz;
    s.c_mm = p.speed_of_sound_m_per_s * 1000.0;
    s.f0 = p.center_frequency_hz;
    // -6 dB full bandwidth B = frac*f0 of a Gaussian spectrum: sigma_f = (B/2)/sqrt(2 ln 2).
    double sigma_f = 0.5 * p.fractional_bandwidth * p.center_frequency_hz / sqrt(2.0 * log(2.0));
    s.sigma_t = 1.0 / (2.0 * M_PI * sigma_f);
    s.half_window = 4.0 * s.sigma_t;
    s.beam2 = 2.0 * p.beam_sigma_mm * p.beam_sigma_mm;
    return s;
}

// Adds the echo of scatterer (x, y, z, amp) to a fixed-point line through ADD(n, q).
template <typename AddFn>
__host__ __device__ inline void add_echo(const PulseShape& ps, const LineGeom& g, const double* sc, uint32_t n_s,
                                         AddFn add) {
    double vx = sc[0] - g.ox, vy = sc[1] - g.oy, vz = sc[2] - g.oz;
    double along = vx * g.dx + vy * g.dy + vz * g.dz;
    if (along <= 0) return;
    double perp2 = vx * vx + vy * vy + vz * vz - along * along;
    if (perp2 < 0) perp2 = 0;
    double beam = exp(-perp2 / ps.beam2);
    if (beam < 1e-6) return;
    double rx = sqrt((sc[0] - g.ex) * (sc[0] - g.ex) + (sc[1] - g.ey) * (sc[1] - g.ey) + (sc[2] - g.ez) * (sc[2] - g.ez));
    double t_arr = g.t0 + (along + rx) / ps.c_mm;
    double lo = ceil((t_arr - ps.half_window) * ps.fs);
    double hi = floor((t_arr + ps.half_window) * ps.fs);
    if (hi < 0 || lo >= n_s) return;
    int64_t n0 = lo < 0 ? 0 : static_cast<int64_t>(lo);
    int64_t n1 = hi >= n_s ? static_cast<int64_t>(n_s) - 1 : static_cast<int64_t>(hi);
    const float amp = static_cast<float>(sc[3] * beam);
    const float inv2s2 = static_cast<float>(1.0 / (2.0 * ps.sigma_t * ps.sigma_t));
    const float w0 = static_cast<float>(2.0 * M_PI * ps.f0);
    for (int64_t n = n0; n <= n1; n++) {
        float t = static_cast<float>(static_cast<double>(n) / ps.fs - t_arr);
        float v = amp * expf(-t * t * inv2s2) * cosf(w0 * t);
        add(static_cast<uint32_t>(n), static_cast<int32_t>(rintf(v * static_cast<float>(kFixedScale))));
    }
}

int validate(const supra_synth_params* p) {
    if (!p) return fail(SUPRA_ERR_INVALID_ARG, "synth: null parameters");
    if (p->num_tx == 0 || p->num_channels == 0 || p->num_samples == 0 || p->num_elements == 0)
        return fail(SUPRA_ERR_INVALID_ARG, "synth: empty shape");
    if (!p->tx_origin || !p->tx_direction || !p->tx_t0 || !p->element_pos ||
        (p->num_scatterers && !p->scatterers))
        return fail(SUPRA_ERR_INVALID_ARG, "synth: null table");
    if (p->num_channels < p->num_elements && (!p->tx_first_element || !p->tx_last_element))
        return fail(SUPRA_ERR_INVALID_ARG, "synth: muxed channels need the per-transmit aperture");
    if (!(p->sampling_frequency_hz > 0) || !(p->speed_of_sound_m_per_s > 0) || !(p->center_frequency_hz > 0) ||
        !(p->fractional_bandwidth > 0) || !(p->peak > 0) || !(p->beam_sigma_mm > 0))
        return fail(SUPRA_ERR_INVALID_ARG, "synth: physical parameters must be positive");
    if (p->num_scatterers > 100000) return fail(SUPRA_ERR_TOO_LARGE, "synth: too many scatterers");
    return SUPRA_OK;
}

// Static block partition of [0, n) over the host's hardware threads.
template <typename Fn>
void parallel_for(int64_t n, Fn fn) {
    unsigned hw = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    int64_t per = (n + hw - 1) / hw;
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < hw; w++) {
        int64_t lo = w * per, hi = std::min<int64_t>(n, lo + per);
        if (lo >= hi) break;
        pool.emplace_back([=]() {
            for (int64_t i = lo; i < hi; i++) fn(i);
        });
    }
    for (auto& th : pool) th.join();
}

double noise_sigma(const supra_synth_params* p, const std::vector<double>& line_sumsq) {
    double total = 0;
    for (double v : line_sumsq) total += v;
    double n = static_cast<double>(p->num_tx) * p->num_channels * p->num_samples;
    double rms = std::sqrt(total / n);
    return rms / std::pow(10.0, p->snr_db / 20.0);
}

// ---- device kernels ----------------------------------------------------------

struct DevTables {
    const double* scat;
    const double* origin;
    const double* dir;
    const double* t0;
    const uint32_t* first;
    const uint32_t* last;
    const double* elem;
};

__global__ __launch_bounds__(256) void echo_kernel(supra_synth_params p, DevTables t, int32_t* lines,
                                                   double* line_sumsq) {
    extern __shared__ int32_t s_line[];
    __shared__ double s_red[256];
    const uint32_t line = blockIdx.x;
    const uint32_t tx = line / p.num_channels, ch = line % p.num_channels;
    for (uint32_t n = threadIdx.x; n < p.num_samples; n += blockDim.x) s_line[n] = 0;
    __syncthreads();
    LineGeom g = line_geometry(p, t.origin, t.dir, t.t0, t.first, t.last, t.elem, tx, ch);
    PulseShape ps = pulse_shape(p);
    if (g.has_element) {
        for (uint32_t j = threadIdx.x; j < p.num_scatterers; j += blockDim.x)
            add_echo(ps, g, t.scat + 4 * j, p.num_samples, [&](uint32_t n, int32_t q) { atomicAdd(&s_line[n], q); });
    }
    __syncthreads();
    double acc = 0;
    int32_t* dst = lines + static_cast<uint64_t>(line) * p.num_samples;
    for (uint32_t n = threadIdx.x; n < p.num_samples; n += blockDim.x) {
        int32_t q = s_line[n];
        dst[n] = q;
        double v = q / kFixedScale;
        acc += v * v;
    }
    s_red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s_red[threadIdx.x] += s_red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) line_sumsq[line] = s_red[0];
}

__global__ __launch_bounds__(256) void maxabs_kernel(const int32_t* lines, uint64_t n, float sigma, uint64_t seed,
                                                     unsigned int* maxbits) {
    float m = 0.0f;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        float v = static_cast<float>(lines[i] / kFixedScale) + sigma * gauss(seed, i);
        m = fmaxf(m, fabsf(v));
    }
    atomicMax(maxbits, __float_as_uint(m));
}

template <typename T>
__global__ __launch_bounds__(256) void finish_kernel(const int32_t* lines, uint64_t n, float sigma, uint64_t seed,
                                                     float gain, T* out) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        float v = (static_cast<float>(lines[i] / kFixedScale) + sigma * gauss(seed, i)) * gain;
        if constexpr (sizeof(T) == 2) out[i] = static_cast<T>(rintf(v));
        else out[i] = v;
    }
}

template <typename T>
void finish_host(const int32_t* lines, uint64_t n, float sigma, uint64_t seed, float gain, T* out) {
    parallel_for(static_cast<int64_t>(n), [=](int64_t i) {
        float v = (static_cast<float>(lines[i] / kFixedScale) + sigma * gauss(seed, i)) * gain;
        if constexpr (sizeof(T) == 2) out[i] = static_cast<T>(rintf(v));
        else out[i] = v;
    });
}

}  // namespace
}  // namespace supra_amd

using namespace supra_amd;

extern "C" double supra_synth_uniform(uint64_t seed, uint64_t index) { return uniform01(seed, index); }

extern "C" int supra_synth_rf_host(const supra_synth_params* p, void* out, int32_t out_dtype) {
    clear_error();
    int st = validate(p);
    if (st != SUPRA_OK) return st;
    if (!out) return fail(SUPRA_ERR_INVALID_ARG, "synth: null output");
    const uint64_t nlines = static_cast<uint64_t>(p->num_tx) * p->num_channels;
    const uint64_t n = nlines * p->num_samples;
    std::vector<int32_t> lines(n, 0);
    std::vector<double> sumsq(nlines, 0.0);
    const PulseShape ps = pulse_shape(*p);
    int32_t* lines_p = lines.data();
    double* sumsq_p = sumsq.data();
    parallel_for(static_cast<int64_t>(nlines), [=](int64_t l) {
        const uint32_t tx = static_cast<uint32_t>(l / p->num_channels), ch = static_cast<uint32_t>(l % p->num_channels);
        LineGeom g = line_geometry(*p, p->tx_origin, p->tx_direction, p->tx_t0, p->tx_first_element,
                                   p->tx_last_element, p->element_pos, tx, ch);
        int32_t* dst = lines_p + l * p->num_samples;
        if (g.has_element) {
            for (uint32_t j = 0; j < p->num_scatterers; j++)
                add_echo(ps, g, p->scatterers + 4 * j, p->num_samples, [&](uint32_t k, int32_t q) { dst[k] += q; });
        }
        double acc = 0;
        for (uint32_t k = 0; k < p->num_samples; k++) {
            double v = dst[k] / kFixedScale;
            acc += v * v;
        }
        sumsq_p[l] = acc;
    });
    const float sigma = static_cast<float>(noise_sigma(p, sumsq));
    float m = 0.0f;
    for (uint64_t i = 0; i < n; i++) {
        float v = static_cast<float>(lines[i] / kFixedScale) + sigma * gauss(p->noise_seed, i);
        m = std::max(m, std::fabs(v));
    }
    const float gain = (m > 0) ? static_cast<float>(p->peak / m) : 1.0f;
    if (out_dtype == SUPRA_INT16) finish_host(lines.data(), n, sigma, p->noise_seed, gain, static_cast<int16_t*>(out));
    else finish_host(lines.data(), n, sigma, p->noise_seed, gain, static_cast<float*>(out));
    return SUPRA_OK;
}

extern "C" int supra_synth_rf_device(const supra_synth_params* p, void* out, int32_t out_dtype, void* stream) {
    clear_error();
    int st = validate(p);
    if (st != SUPRA_OK) return st;
    if (!out) return fail(SUPRA_ERR_INVALID_ARG, "synth: null output");
    if (p->num_samples > kMaxDeviceSamples)
        return fail(SUPRA_ERR_TOO_LARGE, "synth: device path holds at most %u samples per line", kMaxDeviceSamples);
    hipStream_t hs = static_cast<hipStream_t>(stream);
    const uint64_t nlines = static_cast<uint64_t>(p->num_tx) * p->num_channels;
    const uint64_t n = nlines * p->num_samples;
    if (nlines >= (1ull << 31)) return fail(SUPRA_ERR_TOO_LARGE, "synth: too many lines");

    auto size_of = [](uint64_t count, size_t elem) { return static_cast<size_t>(count) * elem; };
    std::vector<void*> allocs;
    auto release = [&]() {
        for (void* q : allocs) (void)hipFree(q);
    };
    auto dev_copy = [&](const void* src, size_t bytes, void** dst) -> hipError_t {
        *dst = nullptr;
        if (!src || bytes == 0) return hipSuccess;
        hipError_t e = hipMalloc(dst, bytes);
        if (e != hipSuccess) return e;
        allocs.push_back(*dst);
        return hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
    };
    DevTables t{};
    void* q;
    hipError_t e = hipSuccess;
    if ((e = dev_copy(p->scatterers, size_of(p->num_scatterers, 4 * sizeof(double)), &q)) != hipSuccess) goto hip_err;
    t.scat = static_cast<const double*>(q);
    if ((e = dev_copy(p->tx_origin, size_of(p->num_tx, 3 * sizeof(double)), &q)) != hipSuccess) goto hip_err;
    t.origin = static_cast<const double*>(q);
    if ((e = dev_copy(p->tx_direction, size_of(p->num_tx, 3 * sizeof(double)), &q)) != hipSuccess) goto hip_err;
    t.dir = static_cast<const double*>(q);
    if ((e = dev_copy(p->tx_t0, size_of(p->num_tx, sizeof(double)), &q)) != hipSuccess) goto hip_err;
    t.t0 = static_cast<const double*>(q);
    if ((e = dev_copy(p->tx_first_element, size_of(p->num_tx, sizeof(uint32_t)), &q)) != hipSuccess) goto hip_err;
    t.first = static_cast<const uint32_t*>(q);
    if ((e = dev_copy(p->tx_last_element, size_of(p->num_tx, sizeof(uint32_t)), &q)) != hipSuccess) goto hip_err;
    t.last = static_cast<const uint32_t*>(q);
    if ((e = dev_copy(p->element_pos, size_of(p->num_elements, 3 * sizeof(double)), &q)) != hipSuccess) goto hip_err;
    t.elem = static_cast<const double*>(q);
    {
        int32_t* lines = nullptr;
        double* d_sumsq = nullptr;
        unsigned int* d_max = nullptr;
        if ((e = hipMalloc(&lines, size_of(n, sizeof(int32_t)))) != hipSuccess) goto hip_err;
        allocs.push_back(lines);
        if ((e = hipMalloc(&d_sumsq, size_of(nlines, sizeof(double)))) != hipSuccess) goto hip_err;
        allocs.push_back(d_sumsq);
        if ((e = hipMalloc(&d_max, sizeof(unsigned int))) != hipSuccess) goto hip_err;
        allocs.push_back(d_max);
        supra_synth_params pd = *p;
        hipLaunchKernelGGL(echo_kernel, dim3(static_cast<uint32_t>(nlines)), dim3(256),
                           p->num_samples * sizeof(int32_t), hs, pd, t, lines, d_sumsq);
        if ((e = hipGetLastError()) != hipSuccess) goto hip_err;
        std::vector<double> sumsq(nlines);
        if ((e = hipMemcpyAsync(sumsq.data(), d_sumsq, nlines * sizeof(double), hipMemcpyDeviceToHost, hs)) !=
            hipSuccess)
            goto hip_err;
        if ((e = hipStreamSynchronize(hs)) != hipSuccess) goto hip_err;
        const float sigma = static_cast<float>(noise_sigma(p, sumsq));
        if ((e = hipMemsetAsync(d_max, 0, sizeof(unsigned int), hs)) != hipSuccess) goto hip_err;
        const uint32_t grid = 4096;
        hipLaunchKernelGGL(maxabs_kernel, dim3(grid), dim3(256), 0, hs, lines, n, sigma, p->noise_seed, d_max);
        unsigned int mbits = 0;
        if ((e = hipMemcpyAsync(&mbits, d_max, sizeof(mbits), hipMemcpyDeviceToHost, hs)) != hipSuccess) goto hip_err;
        if ((e = hipStreamSynchronize(hs)) != hipSuccess) goto hip_err;
        float m;
        std::memcpy(&m, &mbits, sizeof(m));
        const float gain = (m > 0) ? static_cast<float>(p->peak / m) : 1.0f;
        if (out_dtype == SUPRA_INT16)
            hipLaunchKernelGGL(finish_kernel<int16_t>, dim3(grid), dim3(256), 0, hs, lines, n, sigma, p->noise_seed,
                               gain, static_cast<int16_t*>(out));
        else
            hipLaunchKernelGGL(finish_kernel<float>, dim3(grid), dim3(256), 0, hs, lines, n, sigma, p->noise_seed,
                               gain, static_cast<float*>(out));
        if ((e = hipGetLastError()) != hipSuccess) goto hip_err;
        if ((e = hipStreamSynchronize(hs)) != hipSuccess) goto hip_err;
    }
    release();
    return SUPRA_OK;
hip_err:
    release();
    return fail(SUPRA_ERR_HIP, "synth: %s", hipGetErrorString(e));
}
