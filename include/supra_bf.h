/*
 * supra_bf.h — C ABI of the MI355X (gfx950) receive beamformer.
 *
 * This library replaces the CUDA engines behind SUPRA's BeamformingNode,
 * RawDelayNode and BeamformingMVNode.  Every entry point below names the
 * reference interface it stands in for (paths relative to the SUPRA source
 * tree).  The ABI uses plain pointers, sizes and POD structs only: device
 * buffers are passed as `void*` / typed pointers that the caller owns, and
 * HIP streams are passed as `void*` (a `hipStream_t`; NULL = default stream).
 *
 * Conventions shared by all entry points
 *   - Return value: SUPRA_OK (0) on success, a negative SUPRA_ERR_* code
 *     otherwise; supra_last_error() returns a message for the calling thread.
 *   - Launches are asynchronous on `stream`; nothing synchronises unless the
 *     function says so (the reference likewise returns without a sync,
 *     src/SupraLib/Beamformer/RxBeamformerCuda.cu:551-558).
 *   - RF layout [tx][channel][sample], sample fastest
 *     (RxSampleBeamformerDelayAndSum.h:181).  Image layout [depth][scanline],
 *     scanline fastest (RxBeamformerCuda.cu:321).  RawDelayed layout
 *     [scanline][channel][depth], depth fastest (RawDelay.cu:307).
 *   - Lengths in the geometry are millimetres and seconds, exactly as
 *     RxBeamformerParameters holds them; the library converts to "dt-space"
 *     (samples of one-way travel) itself on every call whose sampling
 *     frequency or speed of sound differs from the last one.
 */
#ifndef SUPRA_BF_H
#define SUPRA_BF_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define SUPRA_OK                 0
#define SUPRA_ERR_INVALID_ARG   -1
#define SUPRA_ERR_UNSUPPORTED   -2
#define SUPRA_ERR_HIP           -3
#define SUPRA_ERR_NO_DEVICE     -4
#define SUPRA_ERR_TOO_LARGE     -5

/* ---- element types (USRawData / USImage payloads) ---------------------- */
#define SUPRA_INT16    1   /* int16_t: RF from the probe, or clamped image */
#define SUPRA_FLOAT32  2   /* float */
#define SUPRA_UINT8    3   /* uint8_t: log-compressed B-mode image */

/* ---- WindowType (src/SupraLib/Beamformer/WindowFunction.h:37-44) ------- */
#define SUPRA_WINDOW_RECTANGULAR 0
#define SUPRA_WINDOW_HANN        1
#define SUPRA_WINDOW_HAMMING     2
#define SUPRA_WINDOW_GAUSS       3

/* ---- RxSampleBeamformer (src/SupraLib/Beamformer/RxBeamformerCuda.h:31-37) */
#define SUPRA_BF_DELAY_AND_SUM     0
#define SUPRA_BF_DELAY_AND_STD_DEV 1   /* RxSampleBeamformerDelayAndStdDev */
#define SUPRA_BF_COHERENCE_FACTOR  2   /* RxSampleBeamformerCoherenceFactorDelayAndSum */
#define SUPRA_BF_TEST_SIGNAL       3

/* Number of window-LUT entries (RxBeamformerCuda.h:78). */
#define SUPRA_WINDOW_NUM_ENTRIES 64

/* Marker of an element that is connected to no channel
 * (src/SupraLib/Beamformer/USTransducer.h:35). */
#define SUPRA_ELEMENT_NOT_CONNECTED (-1)

/* ---- ScanlineRxParameters3D (src/SupraLib/USImageProperties.h:43-95) ----
 * Field for field the reference struct, in an explicit C layout.  Positions
 * in mm, initial_delay in seconds.  `first`/`last` are the producer's
 * inclusive aperture corners; the beamformers loop first <= e < last
 * (RxSampleBeamformerDelayAndSum.h:55-57,150). */
typedef struct supra_tx_params {
    uint16_t first_active_element[2];   /* x, y */
    uint16_t last_active_element[2];    /* x, y */
    uint16_t tx_scanline_idx;
    uint16_t reserved[3];
    double   initial_delay;             /* [s] */
} supra_tx_params;

typedef struct supra_scanline_rx {
    double position[3];                 /* x, y(depth axis), z(elevation) [mm] */
    double direction[3];                /* unit vector, depth component in [1] */
    double tx_weights[4];
    supra_tx_params tx[4];
    double max_element_distance[2];     /* [mm] */
} supra_scanline_rx;

/* ---- RxBeamformerParameters (src/SupraLib/Beamformer/RxBeamformerParameters.h:38-112)
 * Host arrays; the library copies them in supra_bf_create. */
typedef struct supra_rx_geometry {
    uint32_t num_rx_scanlines;          /* rx_scanline_layout[0] * [1] */
    uint32_t rx_scanline_layout[2];     /* 3D iff both > 1 (RxBeamformerCuda.cu:31) */
    uint32_t num_depths;
    uint32_t num_elements;
    uint32_t element_layout[2];         /* elements per row / rows */
    const supra_scanline_rx* scanlines; /* num_rx_scanlines, x fastest */
    const float* depths;                /* num_depths [mm] */
    const float* element_x;             /* num_elements [mm] */
    const float* element_y;             /* num_elements [mm] (the "z" of the kernels) */
    const int32_t* element_to_channel;  /* NULL: channel = element % num_channels */
    double speed_of_sound_mm_per_s;     /* the geometry's c */
    double depth;                       /* imaging depth [mm] */
} supra_rx_geometry;

/* ---- Beamformer configuration (src/SupraLib/Beamformer/Beamformer.h, setters
 *      Beamformer.cpp:91-360) and USTransducer (USTransducer.cpp:23-49). ---- */
#define SUPRA_SCAN_LINEAR     0
#define SUPRA_SCAN_PHASED     1
#define SUPRA_SCAN_BIPHASED   2
#define SUPRA_SCAN_PLANEWAVE  3
#define SUPRA_TRANSDUCER_LINEAR 0
#define SUPRA_TRANSDUCER_PLANAR 1

typedef struct supra_beamformer_config {
    int32_t  scan_type;                 /* SUPRA_SCAN_* */
    int32_t  transducer_type;           /* SUPRA_TRANSDUCER_* */
    uint32_t element_layout[2];         /* linear: {N, 1} */
    double   pitch[2];                  /* uniform element pitch [mm] (x, y) */
    uint32_t num_scanlines[2];          /* transmit events (x, y) */
    uint32_t rx_scanline_subdivision[2];
    uint32_t max_aperture[2];           /* receive aperture (elements) */
    uint32_t tx_max_aperture[2];        /* transmit aperture (elements) */
    double   depth;                     /* [mm] */
    uint32_t num_depths;                /* reconstructed samples per scanline */
    double   steering_angle_deg[2];
    double   sector_angle_deg[2];
    int32_t  tx_focus_active;
    double   tx_focus_depth;            /* [mm] */
    double   tx_focus_width;            /* [mm] */
    double   speed_of_sound;            /* [m/s] */
} supra_beamformer_config;

/* Size query for supra_geometry_build: number of rx scanlines and elements. */
int supra_geometry_sizes(const supra_beamformer_config* cfg,
                         uint32_t* num_rx_scanlines, uint32_t rx_layout[2],
                         uint32_t* num_elements);

/* Beamformer::computeTxParameters + getCurrentRxBeamformerParameters
 * (src/SupraLib/Beamformer/Beamformer.cpp:409-838, RxBeamformerParameters.cpp:38-88,
 *  USTransducer::computeInternals USTransducer.cpp:179-258).
 * Fills caller arrays: scanlines[num_rx_scanlines], depths[num_depths],
 * element_x/element_y[num_elements]. */
int supra_geometry_build(const supra_beamformer_config* cfg,
                         supra_scanline_rx* scanlines, float* depths,
                         float* element_x, float* element_y);

/* ---- engine context ------------------------------------------------------ */
typedef struct supra_bf_ctx supra_bf_ctx;

/* RxBeamformerCuda::RxBeamformerCuda(const RxBeamformerParameters&)
 * (src/SupraLib/Beamformer/RxBeamformerCuda.cu:24-52); also stands in for
 * RawDelay::RawDelay (RawDelay.cu:34-60).  Copies the geometry to the
 * current HIP device. */
int  supra_bf_create(const supra_rx_geometry* geometry, supra_bf_ctx** out_ctx);
void supra_bf_destroy(supra_bf_ctx* ctx);

/* RxBeamformerCuda::performRxBeamforming<ChannelDataType, ImageDataType>
 * (src/SupraLib/Beamformer/RxBeamformerCuda.cu:539-663; kernels :109-323).
 * rf:  device [num_tx][num_channels][num_samples] of rf_dtype
 * out: device [num_depths][num_rx_scanlines] of out_dtype (caller-owned)
 * sample_beamformer: SUPRA_BF_DELAY_AND_SUM, SUPRA_BF_DELAY_AND_STD_DEV
 * (RxSampleBeamformerDelayAndStdDev.h:36-270), SUPRA_BF_COHERENCE_FACTOR
 * (RxSampleBeamformerCoherenceFactorDelayAndSum.h:33-258) or
 * SUPRA_BF_TEST_SIGNAL.  In the two variants an aperture element whose delay
 * is out of range contributes the value 0 (the reference reuses the previous
 * element's value there).
 * Linear interpolation of the RF is always on, as in the reference
 * (RxBeamformerCuda.cu:598,623). */
int supra_das_run(supra_bf_ctx* ctx, int32_t sample_beamformer,
                  const void* rf, int32_t rf_dtype,
                  uint32_t num_tx, uint32_t num_channels, uint32_t num_samples,
                  double sampling_frequency_hz,
                  double f_number, double speed_of_sound_mm_per_s,
                  int32_t window_type, float window_parameter,
                  int32_t interpolate_transmits, int32_t additional_offset,
                  void* out, int32_t out_dtype, void* stream);

/* The same as supra_das_run for `num_frames` frames that share geometry and
 * parameters: rf is [frame][tx][ch][s], out [frame][depth][scanline].
 * No reference counterpart (the reference beamforms one USRawData per call,
 * BeamformingNode.cpp:128-177); it exists so that short frames fill the GPU. */
int supra_das_run_frames(supra_bf_ctx* ctx, int32_t sample_beamformer, uint32_t num_frames,
                         const void* rf, int32_t rf_dtype,
                         uint32_t num_tx, uint32_t num_channels, uint32_t num_samples,
                         double sampling_frequency_hz,
                         double f_number, double speed_of_sound_mm_per_s,
                         int32_t window_type, float window_parameter,
                         int32_t interpolate_transmits, int32_t additional_offset,
                         void* out, int32_t out_dtype, void* stream);

/* Differentiable 2D delay-and-sum, forward.  No reference counterpart: the
 * reference beamforms a frame with one speed of sound
 * (RxBeamformerCuda.cu:539-663) and differentiates nothing.  This is the
 * DelayAndSum beamformer of supra_das_run_frames with FP32 RF and image and
 * one speed of sound per frame, read from the DEVICE array
 * speed_of_sound_mm_per_s[num_frames] (positive, finite; the dt-space
 * tables of every frame are built on the GPU, no host round trip).  For each
 * frame f it produces exactly what supra_das_run does at c = c[f].
 * norm_out (device, may be NULL): [frame][slot][scanline][depth] floats,
 * slot count 4 with interpolate_transmits, else 1: the derivative of the
 * pixel with respect to the slot's weighted channel sum (txWeight * n / W),
 * or 0 where the pixel saturates; supra_das_adjoint consumes it. */
int supra_das_forward_c(supra_bf_ctx* ctx, uint32_t num_frames, const float* rf,
                        uint32_t num_tx, uint32_t num_channels, uint32_t num_samples,
                        double sampling_frequency_hz, double f_number,
                        const double* speed_of_sound_mm_per_s,
                        int32_t window_type, float window_parameter,
                        int32_t interpolate_transmits, int32_t additional_offset,
                        float* out, float* norm_out, void* stream);

/* Largest device scratch of supra_das_forward_c_noalloc, in bytes: the
 * per-frame tables of one call, frames * (num_rx_scanlines * 192 +
 * num_depths * 4 + num_elements * 4 + 4) bytes rounded up per part. */
#define SUPRA_NOALLOC_SCRATCH_BYTES (8u << 20)

/* supra_das_forward_c without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as supra_das_forward_c; the per-frame dt-space
 * tables live in a file-scope device array of SUPRA_NOALLOC_SCRATCH_BYTES
 * per GPU (looked up at supra_bf_create, held until the process ends), and
 * calls whose tables need more return SUPRA_ERR_TOO_LARGE.  Every context on
 * one GPU shares that array: calls that overlap on one GPU (different
 * streams or threads without ordering) race on it; issue them in order on
 * one stream.  The context must be 2D (its device geometry is uploaded at
 * supra_bf_create). */
int supra_das_forward_c_noalloc(supra_bf_ctx* ctx, uint32_t num_frames, const float* rf,
                                uint32_t num_tx, uint32_t num_channels, uint32_t num_samples,
                                double sampling_frequency_hz, double f_number,
                                const double* speed_of_sound_mm_per_s,
                                int32_t window_type, float window_parameter,
                                int32_t interpolate_transmits, int32_t additional_offset,
                                float* out, float* norm_out, void* stream);

/* Adjoint of supra_das_forward_c.  No reference counterpart.  Given
 * grad_image = dL/d(out) (device [frame][depth][scanline]) and the forward's
 * norm_out, writes (each may be NULL):
 *   grad_rf [frame][tx][ch][s] (device, every element written) = dL/d(rf);
 *   grad_c  [frame] (device doubles) = dL/dc in 1/(mm/s), through the
 *           delays: dtau/dc = -(tau - t0 - offset)/c, aperture, window and
 *           n/W held fixed (they depend on c only through rounding).
 * Delays are recomputed, not stored.  Arguments other than norm/grad_* must
 * equal the forward call's.  Calls on one context share its scratch: issue
 * them on one stream. */
int supra_das_adjoint(supra_bf_ctx* ctx, uint32_t num_frames, const float* rf,
                      uint32_t num_tx, uint32_t num_channels, uint32_t num_samples,
                      double sampling_frequency_hz, double f_number,
                      const double* speed_of_sound_mm_per_s,
                      int32_t window_type, float window_parameter,
                      int32_t interpolate_transmits, int32_t additional_offset,
                      const float* norm, const float* grad_image,
                      float* grad_rf, double* grad_c, void* stream);

/* RawDelay::performDelay<ChannelDataType, OutputType>
 * (src/SupraLib/Beamformer/RawDelay.cu:555-650; 2D kernel :404-468, per-sample
 * math :221-312; 3D kernel :315-402, per-sample math :126-220).  For each rx
 * scanline's highest-weight transmit, writes every aperture element's
 * delayed, apodised sample times n/W: out[scanline][j][depth] (device,
 * out_dtype).  2D: j = element - first; channels j >= last - first are
 * written as 0.  3D: j = (ex - first.x) + (ey - first.y) * element_layout[0]
 * (the largest j must be below num_channels); slots no aperture element maps
 * to are written as 0, and so are elements outside the F-number aperture or
 * with their delay out of range.  out_sampling_frequency_hz
 * (may be NULL) receives the reference's output rate 1/(depth/c/num_depths)
 * (RawDelay.cu:642). */
int supra_rawdelay_run(supra_bf_ctx* ctx, const void* rf, int32_t rf_dtype,
                       uint32_t num_tx, uint32_t num_channels, uint32_t num_samples,
                       double sampling_frequency_hz, double f_number, double speed_of_sound_mm_per_s,
                       int32_t window_type, float window_parameter,
                       void* out, int32_t out_dtype, double* out_sampling_frequency_hz, void* stream);

/* RxBeamformerMV::performRxBeamforming<ChannelDataType, ImageDataType>
 * (src/SupraLib/Beamformer/RxBeamformerMV.cu:382-614, kernels :79-378, the
 * cuBLAS solve :558-586), with the cublasHandle_t argument dropped and a
 * diagonal-loading factor added (1.0 = the reference: trace(R)/L_t).
 * raw_delayed: device [num_scanlines][num_channels][num_samples] (the output
 * of supra_rawdelay_run); out: device [num_samples][num_scanlines].
 * sub_array_size 0 selects num_channels/2 (BeamformingMVNode default).
 * One fused launch; no host synchronisation. */
int supra_mv_run(const void* raw_delayed, int32_t in_dtype,
                 uint32_t num_scanlines, uint32_t num_channels, uint32_t num_samples,
                 uint32_t sub_array_size, uint32_t temporal_smoothing,
                 double sub_array_scaling_power, int32_t compute_means, double loading_factor,
                 void* out, int32_t out_dtype, void* stream);

/* RxBeamformerMVpcg::performRxBeamforming (RxBeamformerMVpcg.cu:628-800),
 * the BeamformingMVpcgNode engine: supra_mv_run's covariance, smoothing and
 * loading (trace / L_t), with the weights from Jacobi-preconditioned
 * conjugate gradients instead of a direct solve (solveBlockwisePCGJacobi,
 * :450-546): x0 = 1, the residual recomputed every iteration, at most
 * max_iterations_override iterations (0: the sub-array size, :649), stopping
 * once |1 - R~ w| <= convergence_threshold (the node's
 * 10^convergenceThresholdExponent).  Same layouts and types as supra_mv_run. */
int supra_mvpcg_run(const void* raw_delayed, int32_t in_dtype,
                    uint32_t num_scanlines, uint32_t num_channels, uint32_t num_samples,
                    uint32_t sub_array_size, uint32_t temporal_smoothing,
                    uint32_t max_iterations_override, double convergence_threshold,
                    double sub_array_scaling_power, int32_t compute_means,
                    void* out, int32_t out_dtype, void* stream);

/* The pre-delay of supra_rawdelay_run feeding supra_mv_run, fused into one
 * launch: the pre-delayed samples are computed in LDS per workgroup (tile of
 * depths plus the temporal-smoothing halo) and never written to memory.
 * No reference counterpart (the reference chains RawDelayNode and
 * BeamformingMVNode through a device buffer, RawDelay.cu:555-650 then
 * RxBeamformerMV.cu:382-614).  Output [num_depths][num_rx_scanlines],
 * bit-identical to the two-step path with a float intermediate; the
 * minimum-variance input has num_channels channels. */
int supra_mv_das_run(supra_bf_ctx* ctx, const void* rf, int32_t rf_dtype,
                     uint32_t num_tx, uint32_t num_channels, uint32_t num_samples,
                     double sampling_frequency_hz, double f_number, double speed_of_sound_mm_per_s,
                     int32_t window_type, float window_parameter,
                     uint32_t sub_array_size, uint32_t temporal_smoothing,
                     double sub_array_scaling_power, int32_t compute_means, double loading_factor,
                     void* out, int32_t out_dtype, void* stream);

/* ---- batched stream of uneven acquisitions ------------------------------
 * No reference counterpart (the reference runs one node call per frame).
 * A batch is a fixed list of acquisitions, each with its own context
 * (geometry), RF and output buffers, beamformed by delay-and-sum (the
 * arguments of supra_das_run) or by pre-delay + minimum variance (those of
 * supra_mv_das_run).  supra_batch_create validates and plans the batch once
 * (optionally timing tile candidates on the GPU: autotune_tiles != 0),
 * captures the launch in a HIP graph, and keeps the pointers: later launches
 * beamform whatever the RF buffers then hold.  supra_batch_launch replays
 * the graph asynchronously on `stream`; supra_batch_latencies synchronises
 * with the last launch and returns each acquisition's completion time after
 * the launch started (device wall clock, microseconds) and the largest one.
 * DAS items: 2D, RF int16/float32, image int16/float32.  MV items: 2D, RF
 * int16/float32, float32 image.  At most 256 acquisitions per kernel class.
 * Launches of one batch must not overlap (they share its queue counters);
 * the contexts must outlive the batch. */
#define SUPRA_BATCH_DAS 0
#define SUPRA_BATCH_MV  1

typedef struct supra_batch_item {
    supra_bf_ctx* ctx;
    const void* rf;
    int32_t rf_dtype;
    int32_t mode;                       /* SUPRA_BATCH_DAS or SUPRA_BATCH_MV */
    uint32_t num_tx, num_channels, num_samples;
    int32_t window_type;
    double sampling_frequency_hz, f_number, speed_of_sound_mm_per_s;
    float window_parameter;
    int32_t interpolate_transmits;      /* DAS */
    int32_t additional_offset;          /* DAS */
    uint32_t sub_array_size;            /* MV (0: num_channels / 2) */
    uint32_t temporal_smoothing;        /* MV */
    int32_t compute_means;              /* MV */
    double sub_array_scaling_power;     /* MV */
    double loading_factor;              /* MV */
    void* out;                          /* [num_depths][num_rx_scanlines] */
    int32_t out_dtype;
    int32_t reserved;
} supra_batch_item;

typedef struct supra_batch supra_batch;
int  supra_batch_create(const supra_batch_item* items, uint32_t num_items, int32_t autotune_tiles,
                        supra_batch** out_batch);
int  supra_batch_launch(supra_batch* batch, void* stream);
int  supra_batch_latencies(supra_batch* batch, double* frame_us, double* total_us);
int  supra_batch_config(const supra_batch* batch, char* buffer, size_t buffer_size);
void supra_batch_destroy(supra_batch* batch);

/* ---- raw-data mock files ---------------------------------------------------
 * SUPRA's recorded-acquisition format: metadata in ".json"
 * (RxBeamformerParameters::writeMetaDataForMock / readMetaDataForMockJson,
 * src/SupraLib/Beamformer/RxBeamformerParameters.cpp:93-156, :311-450) or in
 * the ASCII ".mock" v1 format (readMetaDataForMockAscii, :187-309; scanline
 * record of USImageProperties.cpp:59-79), and the frames as consecutive int16
 * [tx][channel][sample] blocks of one file (UltrasoundInterfaceRawDataMock.cpp:
 * 49-92, 187-204).  Host memory only. */
typedef struct supra_mock_meta {
    uint32_t num_elements;
    uint32_t element_layout[2];
    uint32_t num_channels;              /* numReceivedChannels */
    uint32_t num_samples;
    uint32_t num_tx;                    /* numTxScanlines */
    uint32_t rx_scanline_layout[2];
    uint32_t num_depths;                /* rxNumDepths */
    double   depth;                     /* [mm] */
    double   sampling_frequency_hz;
    double   speed_of_sound_mm_per_s;
} supra_mock_meta;

/* Sizes and scalars of a .json or .mock metadata file (by suffix). */
int supra_mock_read_meta(const char* path, supra_mock_meta* meta);
/* The receive geometry of the file into caller arrays sized from
 * supra_mock_read_meta: scanlines[rx_scanline_layout[0] * [1]] (x fastest),
 * depths[num_depths] (.json: depth * i / (num_depths - 1), as the reference
 * deduces them), element_x/element_y[num_elements]. */
int supra_mock_read_geometry(const char* path, supra_scanline_rx* scanlines, float* depths,
                             float* element_x, float* element_y);
/* writeMetaDataForMock: the .json metadata (doubles with 17 significant
 * digits, so a read gives back the same values). */
int supra_mock_write_json(const char* path, const supra_mock_meta* meta, const supra_rx_geometry* geometry);
/* The ASCII .mock v1 layout that readMetaDataForMockAscii reads (doubles with
 * 17 significant digits where the reference's writer used 9). */
int supra_mock_write_ascii(const char* path, const supra_mock_meta* meta, const supra_rx_geometry* geometry);
/* Frames in a raw file: file size / (num_tx * num_channels * num_samples * 2). */
int supra_mock_count_frames(const char* raw_path, const supra_mock_meta* meta, uint64_t* num_frames);
/* Frames first .. first + num_frames - 1 into host memory out
 * [frame][tx][channel][sample]. */
int supra_mock_read_frames(const char* raw_path, const supra_mock_meta* meta, uint64_t first_frame,
                           uint32_t num_frames, int16_t* out);

/* ---- after the beamformer: envelope detection and log compression ------- */

/* HilbertFirEnvelope::demodulate<InputType, OutputType>
 * (src/SupraLib/Beamformer/HilbertFirEnvelope.cu:28-104; filter of
 * FirFilterFactory.h:55-200 with FilterTypeHilbertTransformer and
 * FilterWindowHamming; HilbertFirEnvelopeNode.cpp:42: filterLength 1..257,
 * default 65).  image/out: device [num_frames][num_samples][num_scanlines],
 * scanline fastest (num_frames USImages of one size, each filtered on its
 * own; at most 65535 per call).  in_dtype SUPRA_INT16 or SUPRA_FLOAT32, out_dtype
 * SUPRA_INT16 or SUPRA_FLOAT32.  out = clampCast(sqrt(x^2 + (h * x)^2)), the
 * FIR over samples t - N/2 .. t + N/2 inside the image.  For an even N the
 * reference reads one coefficient past its filter; here that tap adds
 * nothing. */
int supra_envelope_run(const void* image, int32_t in_dtype, uint32_t num_scanlines, uint32_t num_samples,
                       uint32_t num_frames, uint32_t filter_length, void* out, int32_t out_dtype, void* stream);

/* LogCompressor::compress<InputType, OutputType>
 * (src/SupraLib/Beamformer/LogCompressor.cu:20-64; LogCompressorNode.cpp:42-45:
 * dynamicRange 1..200 dB (80), gain 0.01..2 (1), inMax 10..50000 (1024)).
 * count values of in_dtype (SUPRA_UINT8, SUPRA_INT16, SUPRA_FLOAT32) to
 * out_dtype (SUPRA_UINT8 or SUPRA_FLOAT32):
 * clampCast(log10(|x| * 10^(dB/20) / inMax + 1) * gain * 255 / log10(10^(dB/20) + 1)),
 * single precision, inMax converted to the input element type first (as the
 * reference's functor argument). */
int supra_logcompress_run(const void* in, int32_t in_dtype, uint64_t count, double dynamic_range_db,
                          double gain, double in_max, void* out, int32_t out_dtype, void* stream);

/* supra_envelope_run (float envelope) followed by supra_logcompress_run (float
 * input) in one launch: HilbertFirEnvelopeNode -> LogCompressorNode without
 * the intermediate image in memory.  No reference counterpart (the reference
 * chains the two nodes through a device buffer).  out_dtype SUPRA_UINT8 or
 * SUPRA_FLOAT32; the result is bit-identical to the two calls. */
int supra_envelope_logcompress_run(const void* image, int32_t in_dtype, uint32_t num_scanlines,
                                   uint32_t num_samples, uint32_t num_frames, uint32_t filter_length,
                                   double dynamic_range_db, double gain, double in_max,
                                   void* out, int32_t out_dtype, void* stream);

/* Window LUT of the reference (WindowFunction.cpp:18-47, WindowFunction.h:122-138),
 * computed on the host: SUPRA_WINDOW_NUM_ENTRIES floats, max-normalised. */
int supra_window_lut(int32_t window_type, float window_parameter, float* lut_out);

/* Last error message of the calling thread ("" if none). */
const char* supra_last_error(void);

/* Library version string. */
const char* supra_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SUPRA_BF_H */
