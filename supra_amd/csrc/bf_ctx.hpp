// The engine context shared by the entry points of the library, and the
// host helpers that build its dt-space tables.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "bf_internal.hpp"
#include "supra_bf.h"

using supra_amd::ScanlineDT;

struct supra_bf_ctx {
    int device = 0;
    uint32_t n_rx = 0, rx_layout[2] = {0, 0}, n_depths = 0, n_el = 0, el_layout[2] = {0, 0};
    bool is3d = false;
    std::vector<supra_scanline_rx> scan_mm;
    std::vector<float> depths_mm, ex_mm, ey_mm;
    std::vector<int32_t> e2c;
    double geom_c = 0, depth_mm = 0;
    uint32_t max_channel_in_map = 0;

    ScanlineDT* d_scans = nullptr;
    float* d_depths = nullptr;
    float* d_ex = nullptr;
    float* d_ey = nullptr;
    int32_t* d_e2c = nullptr;

    std::vector<ScanlineDT> h_scans;
    std::vector<float> h_depths, h_ex, h_ey;
    double cached_dt = 0, cached_c = 0;
    // Recorded after every launch that reads d_scans/d_depths/d_ex/d_ey; a
    // table upload for another (dt, c) waits for it.
    hipEvent_t tables_used = nullptr;
    bool tables_used_recorded = false;
    bool far_ok = false;   // dt-space depths are >= 0 and non-decreasing
    bool far_ok_mm = false;   // the same for the millimetre depths (any c)
    std::mutex mu;

    // Differentiable path: the millimetre geometry on the device, the
    // c-independent part of the scanline records, per-frame tables and
    // scratch, grown on demand.
    supra_scanline_rx* d_scan_mm = nullptr;
    ScanlineDT* d_scan_fixed = nullptr;
    float* d_depths_mm = nullptr;
    float* d_ex_mm = nullptr;
    void* d_frame_tables = nullptr;
    size_t frame_tables_cap = 0;
    void* d_beta = nullptr;
    size_t beta_cap = 0;
    void* d_dc_part = nullptr;
    size_t dc_part_cap = 0;
    // Writers of each (transmit, channel) RF line for the adjoint, cached per
    // (transmit interpolation, num_tx, num_channels).
    int lines_key_interp = -1;
    uint32_t lines_key_tx = 0, lines_key_ch = 0;
    bool lines_exclusive = false;
    uint32_t n_zero_lines = 0;
    uint32_t* d_zero_lines = nullptr;
    uint32_t max_aperture = 0;   // largest last_x - first_x over scanlines and transmits
    // This device's copy of the file-scope scratch array of the
    // allocation-free forward (shared by every context on the device; not owned).
    void* noalloc_scratch = nullptr;

    ~supra_bf_ctx() {
        if (tables_used) (void)hipEventDestroy(tables_used);
        (void)hipFree(d_scans);
        (void)hipFree(d_depths);
        (void)hipFree(d_ex);
        (void)hipFree(d_ey);
        (void)hipFree(d_e2c);
        (void)hipFree(d_scan_mm);
        (void)hipFree(d_scan_fixed);
        (void)hipFree(d_depths_mm);
        (void)hipFree(d_ex_mm);
        (void)hipFree(d_frame_tables);
        (void)hipFree(d_beta);
        (void)hipFree(d_dc_part);
        (void)hipFree(d_zero_lines);
    }
};


namespace supra_amd {

int hip_fail(hipError_t e, const char* what);
// RxBeamformerCuda::convertToDtSpace restated from the millimetre geometry
// into the context's host tables (h_scans, h_depths, h_ex, h_ey).
void build_dt_tables(supra_bf_ctx* ctx, double dt, double c);
int ensure_dt_tables(supra_bf_ctx* ctx, double dt, double c, hipStream_t stream);
int check_device(const supra_bf_ctx* ctx);
int tables_used_on(supra_bf_ctx* ctx, hipStream_t stream);
int check_rawdelay_call(const supra_bf_ctx* ctx, const char* what, int32_t rf_dtype, uint32_t num_tx,
                        uint32_t num_channels, uint32_t num_samples, double fs, double f_number, double c,
                        int32_t window_type, bool allow3d);
DasArgs rawdelay_args(const supra_bf_ctx* ctx, const void* rf, uint32_t num_tx, uint32_t num_channels,
                      uint32_t num_samples, double dt, double f_number, double c, int32_t window_type,
                      float window_parameter);

}  // namespace supra_amd
