// Internal types shared by the host engine and the HIP kernels.
//
// The device-side scanline table is the reference's ScanlineRxParameters3D
// (src/SupraLib/USImageProperties.h:43-95) already converted to dt-space and
// to the single-precision values the kernels read
// (RxBeamformerCuda.cu:58-107 for the conversion, :164-175 and :261-270 for
// the float casts), plus the transmit selection that the reference kernels
// recompute per pixel (:274-303), done once on the host with the same
// double/float comparison.
#pragma once

#include <cstdint>
#include <vector>

#include "supra_bf.h"

namespace supra_amd {

constexpr int kWindowEntries = SUPRA_WINDOW_NUM_ENTRIES;

struct TxDT {
    uint16_t first_x, first_y, last_x, last_y;   // loop is first <= e < last
    uint32_t tx_idx;
    float t0;                                    // initialDelay in samples
    // Far-field threshold on the aperture half-width aDT: at depths with
    // aDT >= far_thr every connected aperture element is inside the F-number
    // aperture and the window argument no longer depends on depth.
    //   2D: max(maxElementDistance.x, max |x_e - x_s|)
    //   3D: max (x_e - x_s)^2 + (z_e - z_s)^2   (aDT is squared in 3D)
    float far_thr;
    uint32_t pad[3];
};

struct ScanlineDT {
    float x, z;                 // position.x, position.z (dt-space)
    float dir_x, dir_y, dir_z;
    float max_ed_x, max_ed_y;   // maxElementDistance (dt-space)
    int32_t sel;                // transmit used without interpolation; -1: none (weight <= 0)
    float txw[4];               // (float) txWeights[k]
    int32_t active_mask;        // bit k: txWeights[k] > 0.0
    int32_t pad[3];
    TxDT tx[4];
};
static_assert(sizeof(TxDT) == 32, "TxDT layout");
static_assert(sizeof(ScanlineDT) == 192, "ScanlineDT layout");

// Everything a DAS launch needs, passed by value as the kernel argument.
struct DasArgs {
    const void* rf;
    void* out;
    const ScanlineDT* scans;
    const float* depths;        // dt-space
    const float* ex;            // element x (dt-space)
    const float* ey;            // element y (dt-space), the kernels' "z"
    const int32_t* e2c;         // element -> channel, or nullptr
    uint64_t rf_frame_stride;   // elements between frames
    uint64_t out_frame_stride;
    uint32_t n_rx, n_depths, n_tx, n_ch, n_s, n_el;
    uint32_t layout_x, layout_y;
    uint32_t n_frames;
    uint32_t depth_tiles;       // ceil(n_depths / tile)
    float F, dt, c;             // fNumber, 1/fs, speed of sound [mm/s] as floats
    float offset;               // (float) additionalOffset
    int32_t max_aperture;       // unused (0)
    int32_t far_ok;             // depth table is non-negative and non-decreasing
    // Per-frame geometry (one speed of sound per frame, 2D): strides between
    // the frames' tables (0: all frames share one table) and the per-frame c
    // as a float (nullptr: every frame uses c).
    uint32_t scan_frame_stride;     // ScanlineDT records
    uint32_t depth_frame_stride;    // floats
    uint32_t el_frame_stride;       // floats
    const float* frame_c;
    // Optional second output of the 2D kernel (float images only): per
    // frame, transmit slot, scanline and depth, d(pixel)/d(channel sum of the
    // slot) = txWeight * n / W, or 0 where the pixel saturates (NaN or
    // beyond FLT_MAX).  Slots: 4 with transmit interpolation, else 1.
    float* norm_out;
    float lut[kWindowEntries];
};

// Per-frame dt-space tables built on the GPU from the millimetre geometry for
// a device array of speeds of sound: the same double-precision expressions
// and float casts as the host-side table builder.
struct FrameTablesArgs {
    const supra_scanline_rx* scan_mm;   // [n_rx]
    const ScanlineDT* scan_fixed;       // [n_rx] c-independent fields (sel, weights, apertures, t0)
    const float* depths_mm;             // [n_d]
    const float* ex_mm;                 // [n_el]
    const int32_t* e2c;                 // or nullptr
    const double* c;                    // [n_frames] mm/s (device)
    ScanlineDT* scans;                  // [frame][n_rx]
    float* depths;                      // [frame][n_d]
    float* ex;                          // [frame][n_el]
    float* frame_c;                     // [frame]
    uint32_t n_rx, n_d, n_el, n_frames;
    double dt;
};

// The adjoint of the 2D DAS with respect to the RF and the speed of sound.
struct AdjointArgs {
    const float* rf;                // [frame][tx][ch][s]
    const float* beta;              // [frame][slot][scanline][depth]: dL/dy * norm
    const ScanlineDT* scans;        // per frame (scan_frame_stride)
    const float* dtab;              // per frame (depth_frame_stride): 4 floats {depth, aDT, 1/aDT, 0} per depth
    const float* ex;                // per frame (el_frame_stride)
    const float* frame_c;           // [frame] float c
    const int32_t* e2c;             // or nullptr
    float* grad_rf;                 // [frame][tx][ch][s] or nullptr
    double* dc_part;                // [frame][blocks_per_frame * waves] partial sums, or nullptr
    uint64_t rf_frame_stride;
    uint32_t scan_frame_stride, depth_frame_stride, el_frame_stride;
    uint32_t n_rx, n_depths, n_ch, n_s, n_slots, egroups, waves;
    float F, dt, offset;
    int32_t exclusive;              // every written line has exactly one writer
    int32_t interp_tx;
    float lut[kWindowEntries];
};

// ---- batched persistent launches ---------------------------------------------

// Shared by the work-queue kernels of one batch launch: the next work item of
// each queue, completed items per frame and each frame's completion time
// (device wall clock), all cleared by the batch's reset kernel.
struct BatchSync {
    uint32_t* head;             // [1] next item of this queue
    uint32_t* done;             // [batch frame] items completed
    uint64_t* stamp;            // [batch frame] completion time
};

struct DasBatchFrame {
    DasArgs a;                  // one frame (n_frames = 1)
    int32_t rf_dtype, out_dtype, interp, map;
    uint32_t n_items;           // work items (scanline x depth tile) of the frame
    uint32_t frame;             // index in the batch
    uint32_t pad[2];
};

struct DasBatchArgs {
    const DasBatchFrame* frames;
    const uint32_t* items;      // [item] = queue frame << 24 | item of the frame
    uint32_t n_items;
    BatchSync sync;
};

// Host helpers (bf_host.cpp).
void build_window_lut(int window_type, float window_parameter, float* lut);
bool scanline_selection(const supra_scanline_rx& s, int32_t* sel, int32_t* active_mask, float txw[4]);

}  // namespace supra_amd
