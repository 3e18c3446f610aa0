// Scan conversion: beamformed scanlines [sample][scanline] resampled onto a
// Cartesian grid, as SUPRA's ScanConverter (src/SupraLib/Beamformer/ScanConverter.cu:430-895,
// ScanConverter.h:25-39) driven by ScanConverterNode (ScanConverterNode.cpp:41-43, 118-200).
//
// Set-up (supra_scanconv_create), once per scan geometry:
//   * bounding box of every scanline's points at depth 0 and at the imaging
//     depth, image size ceil((max - min) / resolution) + 1 per axis
//     (ScanConverter.cu:612-633);
//   * 2D (one row of scanlines): on the host in double precision, for every
//     pair of neighbouring scanlines the pixels of its bounding box inside
//     the two triangles (:674-733) get (t, d) from the iteration over the
//     slerp-interpolated direction (mapToParameters2D :847-895).  Pairs run
//     on several threads; their results are applied in pair order, so a pixel
//     shared by two pairs keeps the later pair's parameters;
//   * 3D (a grid of scanlines): on the GPU in double precision, one launch
//     per quad of neighbouring scanlines over the voxels of its bounding box,
//     five tetrahedra (:77-81) and mapToParameters3D (:183-265); the launches
//     are ordered on one stream, so a voxel shared by two quads keeps the
//     later quad's parameters, as the reference's launch sequence does.
// The tables live in HBM: per pixel a uint32 sample index (0xffffffff
// outside the scan) and two (2D) or three (3D) float weights.
//
// Conversion (supra_scanconv_run): one lane per output pixel, x fastest;
// each lane reads its table entry once and converts `num_frames` frames
// (scanConvert2D :350-381, scanConvert3D :384-428), clampCast to the output
// type.  Arithmetic is the reference's expression, in float, unfused.
//
// Deliberate differences:
//   * the 2D kernel reads the next depth row only where it exists (the 3D
//     kernel's guard, :413); elsewhere the pixel is 0;
//   * 3D parameters are computed in double (the reference: float on the GPU);
//   * SUPRA_SCANCONV_NORMALIZED_DEPTH_WEIGHT makes the 2D depth weight the
//     fraction d / sampleDistance - floor(d / sampleDistance), as in 3D; by
//     default it is d - floor(...) * sampleDistance (a length in mm), as the
//     reference's 2D table stores it (:720-721).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

#include "status.hpp"
#include "supra_bf.h"
#include "supra_scanconv.h"

using namespace supra_amd;

namespace {

constexpr uint32_t kOutside = 0xffffffffu;

struct V3 {
    double x, y, z;
};
__host__ __device__ inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__host__ __device__ inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline double norm3(V3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
__host__ __device__ inline V3 normalize3(V3 a) {
    const double n = norm3(a);
    return {a.x / n, a.y / n, a.z / n};
}
__host__ __device__ inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

constexpr double kEps = 2.2204e-16;              // M_EPS (utilities/utility.h:43)
constexpr int kMaxIterations = 1000;             // ScanConverter.h:28
constexpr double kDistanceThreshold = 1e-5;      // ScanConverter.h:29
constexpr double kTetrahedronThreshold = 1e-9;   // ScanConverter.cu:46

// slerp3 (src/SupraLib/vec.h:424-445)
__host__ __device__ inline V3 slerp3(V3 a, V3 b, double t) {
    const double omega = acos(dot(a, b));
    if (omega < kEps || t < kEps) return a;
    if (1.0 - t < kEps) return b;
    const double s = sin(omega);
    return (sin((1.0 - t) * omega) / s) * a + (sin(t * omega) / s) * b;
}

inline V3 point_line_connection(V3 a, V3 da, V3 x) {
    const V3 conn = x - a;
    return conn - dot(da, conn) * da;
}

// ScanConverter::mapToParameters2D (ScanConverter.cu:847-895): t in [0, 1]
// between the scanlines (a, da) and (b, db), d the distance of x from the
// interpolated scanline's base.
void map_to_parameters_2d(V3 a, V3 b, V3 da, V3 db, V3 x, double* t_out, double* d_out) {
    double low_t = 0, high_t = 1;
    V3 low_conn = point_line_connection(a, da, x);
    V3 high_conn = point_line_connection(b, db, x);
    double low_dist = norm3(low_conn), high_dist = norm3(high_conn);
    if (high_conn.x == 0 && high_conn.y == 0 && high_conn.z == 0) {
        *t_out = high_t;
        *d_out = norm3(x - b);
        return;
    }
    if (low_conn.x == 0 && low_conn.y == 0 && low_conn.z == 0) {
        *t_out = low_t;
        *d_out = norm3(x - a);
        return;
    }
    double dist = 1e10, t = (high_t - low_t) / 2 + low_t;
    V3 base{0, 0, 0};
    for (int it = 0; it < kMaxIterations && dist > kDistanceThreshold; it++) {
        t = (1 - high_dist / (high_dist + low_dist)) * high_t + (1 - low_dist / (high_dist + low_dist)) * low_t;
        base = (1 - t) * a + t * b;
        const V3 conn = point_line_connection(base, slerp3(da, db, t), x);
        dist = norm3(conn);
        if (dot(low_conn, conn) < 0) {
            high_t = t, high_conn = conn, high_dist = dist;
        } else {
            low_t = t, low_conn = conn, low_dist = dist;
        }
    }
    *t_out = t;
    *d_out = norm3(x - base);
}

// barycentricCoordinate2D / pointInsideTriangle (ScanConverter.cu:818-834)
inline double edge_2d(V3 a, V3 b, V3 c) { return (b.x - a.x) * (c.y - a.y) - (b.y - a.y) * (c.x - a.x); }
inline bool inside_triangle(V3 a, V3 b, V3 c, V3 p) {
    return edge_2d(b, c, p) >= 0 && edge_2d(c, a, p) >= 0 && edge_2d(a, b, p) >= 0;
}

// barycentricCoordinate3D (ScanConverter.cu:148-172): det [a 1; b 1; c 1; p 1]
__device__ inline double det4(V3 a, V3 b, V3 c, V3 p) {
    const double axby = a.x * b.y, cypz = c.y * p.z, axbz = a.x * b.z, czpy = c.z * p.y;
    const double aybx = a.y * b.x, cxpz = c.x * p.z, aybz = a.y * b.z, czpx = c.z * p.x;
    const double azbx = a.z * b.x, cxpy = c.x * p.y, azby = a.z * b.y, cypx = c.y * p.x;
    return (axby - aybx) * (c.z - p.z) + (aybz - azby) * (c.x - p.x) + (azbx - axbz) * (c.y - p.y) +
           (cypz - czpy) * (a.x - b.x) - (cxpz - czpx) * (a.y - b.y) + (cxpy - cypx) * (a.z - b.z);
}

// pointInsideTetrahedron (ScanConverter.cu:125-141)
__device__ inline bool inside_tetrahedron(V3 a, V3 b, V3 c, V3 d, V3 p) {
    return det4(a, b, c, d) > 0 && det4(p, b, c, d) >= -kTetrahedronThreshold &&
           det4(a, p, c, d) >= -kTetrahedronThreshold && det4(a, b, p, d) >= -kTetrahedronThreshold &&
           det4(a, b, c, p) >= -kTetrahedronThreshold;
}

__device__ inline V3 point_plane_connection(V3 a, V3 na, V3 x) { return dot(na, x - a) * na; }

// mapToParameters3D (ScanConverter.cu:183-265): (t1, t2) between the four
// scanlines a, ax (x neighbour), ay (y neighbour), axy and the distance d
// along the interpolated scanline; false when x lies outside the quad's
// planes.
__device__ bool map_to_parameters_3d(const V3 p[4], const V3 dir[4], V3 x, double* t1, double* t2, double* d) {
    const V3 a = p[0], ax = p[1], ay = p[2], axy = p[3];
    const V3 da = dir[0], dax = dir[1], day = dir[2], daxy = dir[3];
    const V3 n_xlow = normalize3(cross(da, (ay + day) - a));
    const V3 n_ylow = normalize3(cross((ax + dax) - a, da));
    const V3 n_xhigh = normalize3(cross(dax, (axy + daxy) - ax));
    const V3 n_yhigh = normalize3(cross((axy + daxy) - ay, day));
    double lo_t[2] = {0, 0}, hi_t[2] = {1, 1};
    V3 lo_cx = point_plane_connection(a, n_xlow, x), hi_cx = point_plane_connection(ax, n_xhigh, x);
    V3 lo_cy = point_plane_connection(a, n_ylow, x), hi_cy = point_plane_connection(ay, n_yhigh, x);
    double lo_d[2] = {norm3(lo_cx), norm3(lo_cy)}, hi_d[2] = {norm3(hi_cx), norm3(hi_cy)};
    if (dot(lo_cx, hi_cx) > 0 || dot(lo_cy, hi_cy) > 0) return false;
    double dist[2] = {1e10, 1e10};
    double t[2] = {0.5, 0.5};
    V3 bx1{0, 0, 0}, bx2{0, 0, 0};
    for (int it = 0; it < kMaxIterations && (dist[0] > kDistanceThreshold || dist[1] > kDistanceThreshold); it++) {
        for (int k = 0; k < 2; k++)
            t[k] = (1 - hi_d[k] / (hi_d[k] + lo_d[k])) * hi_t[k] + (1 - lo_d[k] / (hi_d[k] + lo_d[k])) * lo_t[k];
        bx1 = (1 - t[0]) * a + t[0] * ax;
        bx2 = (1 - t[0]) * ay + t[0] * axy;
        const V3 by1 = (1 - t[1]) * a + t[1] * ay;
        const V3 by2 = (1 - t[1]) * ax + t[1] * axy;
        const V3 line = slerp3(slerp3(da, dax, t[0]), slerp3(day, daxy, t[0]), t[1]);
        const V3 nx = normalize3(cross(line, bx2 - bx1));
        const V3 ny = normalize3(cross(by2 - by1, line));
        const V3 cx = point_plane_connection(bx1, nx, x);
        const V3 cy = point_plane_connection(by1, ny, x);
        dist[0] = norm3(cx);
        dist[1] = norm3(cy);
        if (dot(hi_cx, cx) > kEps) hi_t[0] = t[0], hi_cx = cx, hi_d[0] = dist[0];
        else if (dot(lo_cx, cx) > kEps) lo_t[0] = t[0], lo_cx = cx, lo_d[0] = dist[0];
        if (dot(hi_cy, cy) > kEps) hi_t[1] = t[1], hi_cy = cy, hi_d[1] = dist[1];
        else if (dot(lo_cy, cy) > kEps) lo_t[1] = t[1], lo_cy = cy, lo_d[1] = dist[1];
    }
    *t1 = t[0];
    *t2 = t[1];
    *d = norm3(x - ((1 - t[1]) * bx1 + t[1] * bx2));
    return true;
}

struct Quad3D {
    V3 s[4], e[4];          // points at depth 0 and at the imaging depth: (x,y) (x+1,y) (x,y+1) (x+1,y+1)
    V3 pos[4], dir[4];
    int64_t v0[3];          // first voxel of the quad's bounding box
    uint64_t n[3];          // voxels per axis of the bounding box
    uint32_t size[3];
    V3 bb_min;
    double resolution, sample_distance;
    uint32_t base_index;    // scanline index of (x, y)
    uint32_t n_rx;
};

// computeParameterBB3D / computeParametersVoxel3D (ScanConverter.cu:60-120, 277-347)
__global__ __launch_bounds__(256) void tables_3d_kernel(const Quad3D q, uint32_t* __restrict__ idx,
                                                        float* __restrict__ wx, float* __restrict__ wy,
                                                        float* __restrict__ wz, uint64_t n_pix) {
    const uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= q.n[0] * q.n[1] * q.n[2]) return;
    const int64_t vx = q.v0[0] + static_cast<int64_t>(k % q.n[0]);
    const int64_t vy = q.v0[1] + static_cast<int64_t>((k / q.n[0]) % q.n[1]);
    const int64_t vz = q.v0[2] + static_cast<int64_t>(k / (q.n[0] * q.n[1]));
    const V3 p{static_cast<double>(vx) * q.resolution + q.bb_min.x, static_cast<double>(vy) * q.resolution + q.bb_min.y,
               static_cast<double>(vz) * q.resolution + q.bb_min.z};
    const V3 *s = q.s, *e = q.e;
    if (!(inside_tetrahedron(s[0], s[1], s[2], e[0], p) || inside_tetrahedron(s[1], s[3], s[2], e[3], p) ||
          inside_tetrahedron(s[1], e[0], e[1], e[3], p) || inside_tetrahedron(s[2], e[2], e[0], e[3], p) ||
          inside_tetrahedron(s[1], s[2], e[0], e[3], p)))
        return;
    double t1, t2, d;
    if (!map_to_parameters_3d(q.pos, q.dir, p, &t1, &t2, &d)) return;
    const uint64_t o = static_cast<uint64_t>(vx) + static_cast<uint64_t>(vy) * q.size[0] +
                       static_cast<uint64_t>(vz) * q.size[0] * q.size[1];
    if (o >= n_pix) return;
    const double r = d / q.sample_distance;
    const double fl = floor(r);
    idx[o] = static_cast<uint32_t>(fl) * q.n_rx + q.base_index;
    wx[o] = static_cast<float>(t1);
    wy[o] = static_cast<float>(r - fl);
    wz[o] = static_cast<float>(t2);
}

__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t v, uint64_t n) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void mask_from_index_kernel(const uint32_t* __restrict__ idx, uint8_t* __restrict__ m,
                                                              uint64_t n) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) m[i] = idx[i] != kOutside ? 1 : 0;
}

struct ConvArgs {
    const void* in;
    void* out;
    const uint32_t* idx;
    const float* wx;
    const float* wy;
    const float* wz;
    uint64_t n_pixels;
    uint64_t in_frame;      // elements per input frame (n_samples * n_rx)
    uint32_t n_rx_x, n_rx;
    uint32_t num_frames;
};

template <typename T> __device__ inline float in_at(const T* p, uint64_t i) { return static_cast<float>(p[i]); }
// clampCast (utilities/cudaUtility.h:111-147): fmaxf/fminf keep the non-NaN
// operand (NaN lands on the lower limit), then truncation.
template <typename T> __device__ inline T clamp_out(float x);
template <> __device__ inline float clamp_out<float>(float x) { return fminf(fmaxf(x, -FLT_MAX), FLT_MAX); }
template <> __device__ inline int16_t clamp_out<int16_t>(float x) {
    return static_cast<int16_t>(fminf(fmaxf(x, -32767.0f), 32767.0f));
}
template <> __device__ inline uint8_t clamp_out<uint8_t>(float x) {
    return static_cast<uint8_t>(fminf(fmaxf(x, 0.0f), 255.0f));
}

// V pixels per thread, consecutive in the output: V = 4 stores each frame's
// 4 pixels with one packed store; the host takes it when the pixel count is a
// multiple of 4 and the output is aligned to the packed size.
template <typename OutT, int V> struct alignas(V * sizeof(OutT)) OutPack {
    OutT v[V];
};

template <typename InT, typename OutT, bool IS3D, int V>
__global__ __launch_bounds__(256) void scan_convert_kernel(const ConvArgs a) {
    const uint64_t i0 = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * V;
    if (i0 >= a.n_pixels) return;
    const uint64_t n = a.n_rx, nx = a.n_rx_x;
    uint64_t s[V];
    bool valid[V];
    float wx[V], wy[V], wz[V];
    for (int v = 0; v < V; v++) {
        const uint64_t i = i0 + v;
        const uint32_t s32 = a.idx[i];
        s[v] = s32;
        // the last sample row has no successor: the reference's 3D guard, used in 2D too
        valid[v] = s32 != kOutside && (IS3D ? s[v] + 1 + nx + n : s[v] + 1 + n) < a.in_frame;
        wx[v] = valid[v] ? a.wx[i] : 0.0f;
        wy[v] = valid[v] ? a.wy[i] : 0.0f;
        wz[v] = (IS3D && valid[v]) ? a.wz[i] : 0.0f;
    }
    const InT* in = static_cast<const InT*>(a.in);
    OutT* out = static_cast<OutT*>(a.out);
    for (uint32_t f = 0; f < a.num_frames; f++, in += a.in_frame, out += a.n_pixels) {
        OutPack<OutT, V> o;
        for (int v = 0; v < V; v++) {
            float val = 0.0f;
            if (valid[v]) {
                const uint64_t q = s[v];
                if (!IS3D) {
                    val = (1 - wy[v]) * ((1 - wx[v]) * in_at(in, q) + wx[v] * in_at(in, q + 1)) +
                          wy[v] * ((1 - wx[v]) * in_at(in, q + n) + wx[v] * in_at(in, q + 1 + n));
                } else {
                    val = (1 - wy[v]) * ((1 - wz[v]) * ((1 - wx[v]) * in_at(in, q) + wx[v] * in_at(in, q + 1)) +
                                         wz[v] * ((1 - wx[v]) * in_at(in, q + nx) + wx[v] * in_at(in, q + 1 + nx))) +
                          wy[v] * ((1 - wz[v]) * ((1 - wx[v]) * in_at(in, q + n) + wx[v] * in_at(in, q + 1 + n)) +
                                   wz[v] * ((1 - wx[v]) * in_at(in, q + nx + n) +
                                            wx[v] * in_at(in, q + 1 + nx + n)));
                }
            }
            o.v[v] = clamp_out<OutT>(val);
        }
        *reinterpret_cast<OutPack<OutT, V>*>(out + i0) = o;
    }
}

// 2D with 4 pixels per thread (the packed case), input staged in LDS.  A
// block's 1024 consecutive output pixels read a narrow range of input
// elements (a few sample rows of a linear scan); the block finds that range
// once and copies it per frame, converted to float as the direct kernel
// converts each sample, into one of two LDS windows (one barrier per frame;
// the next frame's window is loaded into registers meanwhile).  The window
// holds floats so that neighbouring samples are read at 4-byte aligned LDS
// addresses.  The four neighbours of each pixel are the elements the direct
// kernel reads, so the results are identical.  A block whose range exceeds
// the window (wide sector rows) reads the input directly.
constexpr uint32_t kScWin = 1024;            // elements per window; two windows per block
constexpr int kScPer = kScWin / 256;         // window elements per thread

template <typename InT, typename OutT>
__global__ __launch_bounds__(256) void scan_convert_lds_kernel(const ConvArgs a) {
    constexpr int V = 4;
    __shared__ float win[2][kScWin];
    __shared__ uint64_t red_lo[4], red_hi[4];
    const uint64_t i0 = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * V;
    const bool live = i0 < a.n_pixels;   // n_pixels is a multiple of V
    const uint64_t n = a.n_rx;
    uint64_t s[V];
    bool valid[V];
    float wx[V], wy[V];
    uint64_t lo = ~0ull, hi = 0;
    for (int v = 0; v < V; v++) {
        const uint64_t i = i0 + v;
        const uint32_t s32 = live ? a.idx[i] : kOutside;
        s[v] = s32;
        // the last sample row has no successor: the reference's 3D guard, used in 2D too
        valid[v] = s32 != kOutside && s[v] + 1 + n < a.in_frame;
        wx[v] = valid[v] ? a.wx[i] : 0.0f;
        wy[v] = valid[v] ? a.wy[i] : 0.0f;
        if (valid[v]) {
            lo = min(lo, s[v]);
            hi = max(hi, s[v] + 1 + n);
        }
    }
    // block range [lo, hi] of the elements read
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, static_cast<uint64_t>(__shfl_xor(static_cast<long long>(lo), o, 64)));
        hi = max(hi, static_cast<uint64_t>(__shfl_xor(static_cast<long long>(hi), o, 64)));
    }
    const uint32_t wave = threadIdx.x / 64;
    if (threadIdx.x % 64 == 0) {
        red_lo[wave] = lo;
        red_hi[wave] = hi;
    }
    __syncthreads();
    lo = min(min(red_lo[0], red_lo[1]), min(red_lo[2], red_lo[3]));
    hi = max(max(red_hi[0], red_hi[1]), max(red_hi[2], red_hi[3]));
    const bool staged = lo <= hi && hi - lo < kScWin;
    const uint32_t extent = staged ? static_cast<uint32_t>(hi - lo + 1) : 0u;

    const InT* in = static_cast<const InT*>(a.in);
    OutT* out = static_cast<OutT*>(a.out);
    float pre[kScPer];
    auto fetch = [&](const InT* src) {
        for (int j = 0; j < kScPer; j++) {
            const uint32_t k = threadIdx.x + static_cast<uint32_t>(j) * 256u;
            pre[j] = k < extent ? static_cast<float>(src[lo + k]) : 0.0f;
        }
    };
    if (staged) fetch(in);
    for (uint32_t f = 0; f < a.num_frames; f++, in += a.in_frame, out += a.n_pixels) {
        const uint32_t b = f & 1u;
        if (staged) {
            for (int j = 0; j < kScPer; j++) {
                const uint32_t k = threadIdx.x + static_cast<uint32_t>(j) * 256u;
                if (k < extent) win[b][k] = pre[j];
            }
            __syncthreads();
            if (f + 1 < a.num_frames) fetch(in + a.in_frame);
        }
        OutPack<OutT, V> o;
        for (int v = 0; v < V; v++) {
            float val = 0.0f;
            if (valid[v]) {
                if (staged) {
                    const uint32_t q = static_cast<uint32_t>(s[v] - lo);
                    val = (1 - wy[v]) * ((1 - wx[v]) * win[b][q] + wx[v] * win[b][q + 1]) +
                          wy[v] * ((1 - wx[v]) * win[b][q + n] + wx[v] * win[b][q + 1 + n]);
                } else {
                    const uint64_t q = s[v];
                    val = (1 - wy[v]) * ((1 - wx[v]) * in_at(in, q) + wx[v] * in_at(in, q + 1)) +
                          wy[v] * ((1 - wx[v]) * in_at(in, q + n) + wx[v] * in_at(in, q + 1 + n));
                }
            }
            o.v[v] = clamp_out<OutT>(val);
        }
        if (live) *reinterpret_cast<OutPack<OutT, V>*>(out + i0) = o;
    }
}

V3 point_at(const supra_scanline_rx& sl, double depth) {
    return V3{sl.position[0] + depth * sl.direction[0], sl.position[1] + depth * sl.direction[1],
              sl.position[2] + depth * sl.direction[2]};
}
V3 pos_of(const supra_scanline_rx& sl) { return {sl.position[0], sl.position[1], sl.position[2]}; }
V3 dir_of(const supra_scanline_rx& sl) { return {sl.direction[0], sl.direction[1], sl.direction[2]}; }

int64_t cell(double v, double mn, double res, bool up) {
    const double q = (v - mn) / res;
    return static_cast<int64_t>(up ? std::ceil(q) : std::floor(q));
}

struct Entry2D {
    uint64_t pixel;
    uint32_t idx;
    float wx, wy;
};

unsigned host_threads() {
    const unsigned hc = std::thread::hardware_concurrency();
    return std::max(1u, std::min(16u, hc ? hc : 1u));
}

}  // namespace

struct supra_scanconv {
    int device = 0;
    bool is3d = false;
    uint32_t layout[2] = {0, 0};
    uint32_t n_samples = 0;
    uint32_t size[3] = {1, 1, 1};
    double bb_min[3] = {0, 0, 0}, bb_max[3] = {0, 0, 0}, resolution = 0;
    uint32_t* d_idx = nullptr;
    float* d_w = nullptr;   // wx, wy, wz, n_pixels each
    ~supra_scanconv() {
        (void)hipFree(d_idx);
        (void)hipFree(d_w);
    }
    uint64_t pixels() const { return static_cast<uint64_t>(size[0]) * size[1] * size[2]; }
};

extern "C" int supra_scanconv_create(const supra_scanline_rx* scanlines, const uint32_t rx_layout[2], double depth,
                                     uint32_t num_samples, double image_resolution, uint32_t flags,
                                     supra_scanconv** out) {
    clear_error();
    if (!scanlines || !rx_layout || !out) return fail(SUPRA_ERR_INVALID_ARG, "scanconv: null argument");
    *out = nullptr;
    const uint32_t lx = rx_layout[0], ly = rx_layout[1];
    if (lx == 0 || ly == 0 || num_samples < 2 || !(depth > 0) || !std::isfinite(depth))
        return fail(SUPRA_ERR_INVALID_ARG, "scanconv: empty scan (layout %u x %u, %u samples, depth %g)", lx, ly,
                    num_samples, depth);
    if (flags & ~static_cast<uint32_t>(SUPRA_SCANCONV_NORMALIZED_DEPTH_WEIGHT))
        return fail(SUPRA_ERR_INVALID_ARG, "scanconv: unknown flags 0x%x", flags);
    const bool is3d = !(lx == 1 || ly == 1);   // USImageProperties::is2D (USImageProperties.cpp:342-345)
    if (!is3d && lx < 2)
        return fail(SUPRA_ERR_INVALID_ARG, "scanconv: a 2D scan needs its scanlines along x (layout %u x %u)", lx, ly);
    if (static_cast<uint64_t>(num_samples) * lx * ly >= kOutside)
        return fail(SUPRA_ERR_TOO_LARGE, "scanconv: %u samples x %u scanlines exceed 32-bit indices", num_samples,
                    lx * ly);
    const double sample_distance = depth / (num_samples - 1);                       // getSampleDistance (:326-329)
    const double res = image_resolution > 0 ? image_resolution : sample_distance;   // getImageResolution (:331-340)
    if (!std::isfinite(res)) return fail(SUPRA_ERR_INVALID_ARG, "scanconv: resolution %g", res);
    auto sl = [&](uint32_t x, uint32_t y) -> const supra_scanline_rx& { return scanlines[x + y * lx]; };
    V3 lo{DBL_MAX, DBL_MAX, DBL_MAX}, hi{-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (uint32_t y = 0; y < ly; y++)
        for (uint32_t x = 0; x < lx; x++)
            for (double dd : {0.0, depth}) {
                const V3 p = point_at(sl(x, y), dd);
                lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
                hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
            }
    const double ext[3] = {hi.x - lo.x, hi.y - lo.y, hi.z - lo.z};
    uint32_t size[3];
    uint64_t n_pix = 1;
    for (int k = 0; k < 3; k++) {
        const double cells = std::ceil(ext[k] / res) + 1;
        if (!(cells >= 1) || cells > static_cast<double>(1u << 24))
            return fail(SUPRA_ERR_TOO_LARGE, "scanconv: image axis %d would hold %g pixels", k, cells);
        size[k] = static_cast<uint32_t>(cells);
        n_pix *= size[k];
    }
    if (n_pix > (1ull << 32))
        return fail(SUPRA_ERR_TOO_LARGE, "scanconv: %llu pixels", static_cast<unsigned long long>(n_pix));

    supra_scanconv* sc = new (std::nothrow) supra_scanconv();
    if (!sc) return fail(SUPRA_ERR_INVALID_ARG, "scanconv: out of host memory");
    sc->is3d = is3d;
    sc->layout[0] = lx, sc->layout[1] = ly;
    sc->n_samples = num_samples;
    sc->resolution = res;
    std::copy(size, size + 3, sc->size);
    sc->bb_min[0] = lo.x, sc->bb_min[1] = lo.y, sc->bb_min[2] = lo.z;
    sc->bb_max[0] = hi.x, sc->bb_max[1] = hi.y, sc->bb_max[2] = hi.z;
    const uint32_t n_rx = lx * ly;
    hipError_t e = hipGetDevice(&sc->device);
    if (e == hipSuccess) e = hipMalloc(&sc->d_idx, n_pix * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&sc->d_w, 3 * n_pix * sizeof(float));
    if (e != hipSuccess) {
        delete sc;
        return fail(SUPRA_ERR_HIP, "scanconv: table allocation (%llu pixels): %s",
                    static_cast<unsigned long long>(n_pix), hipGetErrorString(e));
    }

    if (!is3d) {
        // pairs of neighbouring scanlines in the x-y plane, on host threads,
        // applied in pair order afterwards (ScanConverter.cu:674-733)
        const bool normalized = flags & SUPRA_SCANCONV_NORMALIZED_DEPTH_WEIGHT;
        const uint32_t n_pairs = (lx - 1) * ly;
        std::vector<std::vector<Entry2D>> found(n_pairs);
        std::atomic<uint32_t> next{0};
        auto work = [&]() {
            for (uint32_t pr; (pr = next.fetch_add(1)) < n_pairs;) {
                const uint32_t x = pr % (lx - 1), y = pr / (lx - 1);
                const supra_scanline_rx &l0 = sl(x, y), &l1 = sl(x + 1, y);
                const V3 s0 = point_at(l0, 0), s1 = point_at(l1, 0), e0 = point_at(l0, depth), e1 = point_at(l1, depth);
                const double mnx = std::min(std::min(std::min(s0.x, s1.x), e0.x), e1.x);
                const double mny = std::min(std::min(std::min(s0.y, s1.y), e0.y), e1.y);
                const double mxx = std::max(std::max(std::max(s0.x, s1.x), e0.x), e1.x);
                const double mxy = std::max(std::max(std::max(s0.y, s1.y), e0.y), e1.y);
                const int64_t px0 = std::max<int64_t>(0, cell(mnx, lo.x, res, false));
                const int64_t py0 = std::max<int64_t>(0, cell(mny, lo.y, res, false));
                const int64_t px1 = std::min<int64_t>(size[0] - 1, cell(mxx, lo.x, res, true));
                const int64_t py1 = std::min<int64_t>(size[1] - 1, cell(mxy, lo.y, res, true));
                std::vector<Entry2D>& list = found[pr];
                for (int64_t px = px0; px <= px1; px++) {
                    for (int64_t py = py0; py <= py1; py++) {
                        const V3 p{static_cast<double>(px) * res + lo.x, static_cast<double>(py) * res + lo.y, 0.0};
                        if (!(inside_triangle(e1, e0, s0, p) || inside_triangle(s0, s1, e1, p))) continue;
                        double t, d;
                        map_to_parameters_2d(pos_of(l0), pos_of(l1), dir_of(l0), dir_of(l1), p, &t, &d);
                        const double fl = std::floor(d / sample_distance);
                        const double wy = normalized ? d / sample_distance - fl : d - fl * sample_distance;
                        const uint64_t row = static_cast<uint64_t>(fl);
                        list.push_back({static_cast<uint64_t>(px) + static_cast<uint64_t>(py) * size[0],
                                        static_cast<uint32_t>(row * n_rx + x + y * lx), static_cast<float>(t),
                                        static_cast<float>(wy)});
                    }
                }
            }
        };
        std::vector<std::thread> pool;
        const unsigned nt = std::min<unsigned>(host_threads(), n_pairs);
        for (unsigned w = 1; w < nt; w++) pool.emplace_back(work);
        work();
        for (auto& th : pool) th.join();
        std::vector<uint32_t> idx(n_pix, kOutside);
        std::vector<float> w(3 * n_pix, 0.0f);
        for (const auto& list : found)
            for (const Entry2D& en : list) {
                idx[en.pixel] = en.idx;
                w[en.pixel] = en.wx;
                w[n_pix + en.pixel] = en.wy;
            }
        e = hipMemcpy(sc->d_idx, idx.data(), n_pix * sizeof(uint32_t), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(sc->d_w, w.data(), 3 * n_pix * sizeof(float), hipMemcpyHostToDevice);
    } else {
        // quads of neighbouring scanlines, launched in order on one stream
        // (ScanConverter.cu:745-830)
        hipStream_t s = nullptr;
        e = hipStreamCreate(&s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(fill_u32_kernel, dim3(static_cast<unsigned>((n_pix + 255) / 256)), dim3(256), 0, s,
                               sc->d_idx, kOutside, n_pix);
            hipLaunchKernelGGL(fill_u32_kernel, dim3(static_cast<unsigned>((3 * n_pix + 255) / 256)), dim3(256), 0, s,
                               reinterpret_cast<uint32_t*>(sc->d_w), 0u, 3 * n_pix);
            e = hipGetLastError();
        }
        for (uint32_t y = 0; e == hipSuccess && y + 1 < ly; y++) {
            for (uint32_t x = 0; e == hipSuccess && x + 1 < lx; x++) {
                Quad3D q;
                std::memset(&q, 0, sizeof(q));
                const supra_scanline_rx* corner[4] = {&sl(x, y), &sl(x + 1, y), &sl(x, y + 1), &sl(x + 1, y + 1)};
                V3 mn{DBL_MAX, DBL_MAX, DBL_MAX}, mx{-DBL_MAX, -DBL_MAX, -DBL_MAX};
                for (int c = 0; c < 4; c++) {
                    q.s[c] = point_at(*corner[c], 0);
                    q.e[c] = point_at(*corner[c], depth);
                    q.pos[c] = pos_of(*corner[c]);
                    q.dir[c] = dir_of(*corner[c]);
                    for (const V3& v : {q.s[c], q.e[c]}) {
                        mn = {std::min(mn.x, v.x), std::min(mn.y, v.y), std::min(mn.z, v.z)};
                        mx = {std::max(mx.x, v.x), std::max(mx.y, v.y), std::max(mx.z, v.z)};
                    }
                }
                const double mnv[3] = {mn.x, mn.y, mn.z}, mxv[3] = {mx.x, mx.y, mx.z}, lov[3] = {lo.x, lo.y, lo.z};
                bool empty = false;
                for (int k = 0; k < 3; k++) {
                    const int64_t a0 = std::max<int64_t>(0, cell(mnv[k], lov[k], res, false));
                    const int64_t a1 = std::min<int64_t>(size[k] - 1, cell(mxv[k], lov[k], res, true));
                    empty |= a1 < a0;
                    q.v0[k] = a0;
                    q.n[k] = a1 >= a0 ? static_cast<uint64_t>(a1 - a0 + 1) : 0;
                }
                if (empty) continue;
                std::copy(size, size + 3, q.size);
                q.bb_min = lo;
                q.resolution = res;
                q.sample_distance = sample_distance;
                q.base_index = x + y * lx;
                q.n_rx = n_rx;
                const uint64_t work = q.n[0] * q.n[1] * q.n[2];
                hipLaunchKernelGGL(tables_3d_kernel, dim3(static_cast<unsigned>((work + 255) / 256)), dim3(256), 0, s,
                                   q, sc->d_idx, sc->d_w, sc->d_w + n_pix, sc->d_w + 2 * n_pix, n_pix);
                e = hipGetLastError();
            }
        }
        if (s) {
            const hipError_t e2 = hipStreamSynchronize(s);
            if (e == hipSuccess) e = e2;
            (void)hipStreamDestroy(s);
        }
    }
    if (e != hipSuccess) {
        delete sc;
        return fail(SUPRA_ERR_HIP, "scanconv: tables: %s", hipGetErrorString(e));
    }
    *out = sc;
    return SUPRA_OK;
}

extern "C" void supra_scanconv_destroy(supra_scanconv* sc) { delete sc; }

extern "C" int supra_scanconv_geometry(const supra_scanconv* sc, uint32_t size[3], double bb_min[3],
                                       double bb_max[3], double* resolution) {
    clear_error();
    if (!sc) return fail(SUPRA_ERR_INVALID_ARG, "scanconv: null argument");
    if (size) std::copy(sc->size, sc->size + 3, size);
    if (bb_min) std::copy(sc->bb_min, sc->bb_min + 3, bb_min);
    if (bb_max) std::copy(sc->bb_max, sc->bb_max + 3, bb_max);
    if (resolution) *resolution = sc->resolution;
    return SUPRA_OK;
}

extern "C" int supra_scanconv_tables(const supra_scanconv* sc, uint8_t* mask, uint32_t* sample_index,
                                     float* weights) {
    clear_error();
    if (!sc) return fail(SUPRA_ERR_INVALID_ARG, "scanconv: null argument");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != sc->device)
        return fail(SUPRA_ERR_INVALID_ARG, "scanconv: created on device %d, current device %d", sc->device, dev);
    const uint64_t n = sc->pixels();
    std::vector<uint32_t> idx(n);
    hipError_t e = hipMemcpy(idx.data(), sc->d_idx, n * sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess && weights) e = hipMemcpy(weights, sc->d_w, 3 * n * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(SUPRA_ERR_HIP, "scanconv: table download: %s", hipGetErrorString(e));
    for (uint64_t i = 0; i < n; i++) {
        if (mask) mask[i] = idx[i] != kOutside ? 1 : 0;
        if (sample_index) sample_index[i] = idx[i];
    }
    return SUPRA_OK;
}

extern "C" int supra_scanconv_mask(const supra_scanconv* sc, uint8_t* mask_device, void* stream) {
    clear_error();
    if (!sc || !mask_device) return fail(SUPRA_ERR_INVALID_ARG, "scanconv: null argument");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != sc->device)
        return fail(SUPRA_ERR_INVALID_ARG, "scanconv: created on device %d, current device %d", sc->device, dev);
    const uint64_t n = sc->pixels();
    hipLaunchKernelGGL(mask_from_index_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), sc->d_idx, mask_device, n);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SUPRA_ERR_HIP, "scanconv: mask: %s", hipGetErrorString(e));
    return SUPRA_OK;
}

extern "C" int supra_scanconv_run(const supra_scanconv* sc, const void* scanlines_in, int32_t in_dtype,
                                  uint32_t num_frames, void* out, int32_t out_dtype, void* stream) {
    clear_error();
    if (!sc || !scanlines_in || !out) return fail(SUPRA_ERR_INVALID_ARG, "scanconv: null argument");
    const bool in_ok = in_dtype == SUPRA_UINT8 || in_dtype == SUPRA_INT16 || in_dtype == SUPRA_FLOAT32;
    const bool out_ok = out_dtype == SUPRA_UINT8 || out_dtype == SUPRA_INT16 || out_dtype == SUPRA_FLOAT32;
    if (!in_ok || !out_ok) return fail(SUPRA_ERR_UNSUPPORTED, "scanconv: element types %d -> %d", in_dtype, out_dtype);
    if (num_frames == 0) return SUPRA_OK;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != sc->device)
        return fail(SUPRA_ERR_INVALID_ARG, "scanconv: created on device %d, current device %d", sc->device, dev);
    ConvArgs a;
    std::memset(&a, 0, sizeof(a));
    a.in = scanlines_in;
    a.out = out;
    a.idx = sc->d_idx;
    a.n_pixels = sc->pixels();
    a.wx = sc->d_w;
    a.wy = sc->d_w + a.n_pixels;
    a.wz = sc->d_w + 2 * a.n_pixels;
    a.n_rx_x = sc->layout[0];
    a.n_rx = sc->layout[0] * sc->layout[1];
    a.in_frame = static_cast<uint64_t>(sc->n_samples) * a.n_rx;
    a.num_frames = num_frames;
    const size_t out_size = out_dtype == SUPRA_FLOAT32 ? 4 : (out_dtype == SUPRA_INT16 ? 2 : 1);
    const bool packed = a.n_pixels % 4 == 0 && reinterpret_cast<uintptr_t>(out) % (4 * out_size) == 0;
    const uint64_t threads = packed ? a.n_pixels / 4 : a.n_pixels;
    const dim3 grid(static_cast<uint32_t>((threads + 255) / 256)), block(256);
    hipStream_t s = static_cast<hipStream_t>(stream);
#define SUPRA_SC_LAUNCH(IN, OUT)                                                                                       \
    do {                                                                                                              \
        if (packed) {                                                                                                 \
            if (sc->is3d) hipLaunchKernelGGL((scan_convert_kernel<IN, OUT, true, 4>), grid, block, 0, s, a);          \
            else hipLaunchKernelGGL((scan_convert_lds_kernel<IN, OUT>), grid, block, 0, s, a);                        \
        } else {                                                                                                      \
            if (sc->is3d) hipLaunchKernelGGL((scan_convert_kernel<IN, OUT, true, 1>), grid, block, 0, s, a);          \
            else hipLaunchKernelGGL((scan_convert_kernel<IN, OUT, false, 1>), grid, block, 0, s, a);                  \
        }                                                                                                             \
    } while (0)
#define SUPRA_SC_OUT(IN)                                                                                              \
    do {                                                                                                              \
        if (out_dtype == SUPRA_FLOAT32) SUPRA_SC_LAUNCH(IN, float);                                                   \
        else if (out_dtype == SUPRA_INT16) SUPRA_SC_LAUNCH(IN, int16_t);                                              \
        else SUPRA_SC_LAUNCH(IN, uint8_t);                                                                            \
    } while (0)
    if (in_dtype == SUPRA_FLOAT32) SUPRA_SC_OUT(float);
    else if (in_dtype == SUPRA_INT16) SUPRA_SC_OUT(int16_t);
    else SUPRA_SC_OUT(uint8_t);
#undef SUPRA_SC_OUT
#undef SUPRA_SC_LAUNCH
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SUPRA_ERR_HIP, "scanconv: launch: %s", hipGetErrorString(e));
    return SUPRA_OK;
}
