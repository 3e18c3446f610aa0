/*
 * supra_scanconv.h — scan conversion of beamformed scanlines onto a
 * Cartesian pixel grid, replacing SUPRA's ScanConverter
 * (src/SupraLib/Beamformer/ScanConverter.h:25-93, ScanConverter.cu:430-895)
 * as driven by ScanConverterNode (ScanConverterNode.cpp:118-200).
 *
 * Input: a USImage of the beamformer, [frame][sample][scanline] with the
 * scanline index fastest (x + y * layout[0]), element type uint8, int16 or
 * float.  Output: [frame][z][y][x] with x fastest, element type uint8, int16
 * or float (clampCast).  The geometry (scanline positions and directions,
 * the same array supra_bf_create takes) fixes the pixel tables, built once
 * by supra_scanconv_create and kept in HBM.
 *
 * Conventions as elsewhere in libsupra_bf.so: SUPRA_OK or a SUPRA_ERR_* code
 * (message in supra_last_error()), device pointers, `stream` a hipStream_t
 * (NULL: the null stream), launches asynchronous.
 */
#ifndef SUPRA_SCANCONV_H
#define SUPRA_SCANCONV_H

#include <stdint.h>

#include "supra_bf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* 2D only: depth weight = d / sampleDistance - floor(d / sampleDistance)
 * (a fraction, as the reference's 3D tables hold it).  Without the flag the
 * 2D depth weight is d - floor(d / sampleDistance) * sampleDistance, as the
 * reference's 2D tables hold it (ScanConverter.cu:720-721). */
#define SUPRA_SCANCONV_NORMALIZED_DEPTH_WEIGHT 1u

typedef struct supra_scanconv supra_scanconv;

/* ScanConverter::updateInternals (ScanConverter.cu:516-810).  `scanlines`:
 * rx_layout[0] * rx_layout[1] entries, x fastest (only position and
 * direction are read); 2D when rx_layout[0] == 1 or rx_layout[1] == 1
 * (USImageProperties::is2D, USImageProperties.cpp:342-345), and a 2D scan
 * must run along x.  depth: imaging depth [mm]; num_samples: samples per
 * scanline of the input image (sample distance depth / (num_samples - 1),
 * USImageProperties.cpp:326-329); image_resolution: pixel pitch [mm], <= 0
 * for the sample distance (getImageResolution, :331-340; the node's
 * "imageResolution"/"imageResolutionForced", ScanConverterNode.cpp:41-42,
 * 162-176).  Builds the tables on the current device (2D on host threads, 3D
 * with GPU kernels) and returns when they are resident. */
int supra_scanconv_create(const supra_scanline_rx* scanlines, const uint32_t rx_layout[2], double depth,
                          uint32_t num_samples, double image_resolution, uint32_t flags, supra_scanconv** out);

void supra_scanconv_destroy(supra_scanconv* sc);

/* ScanConverter::getImageSize and the bounding box (ScanConverter.cu:612-633):
 * size[3] pixels along x, y, z; bb_min/bb_max [mm]; resolution [mm].  Any
 * output may be NULL. */
int supra_scanconv_geometry(const supra_scanconv* sc, uint32_t size[3], double bb_min[3], double bb_max[3],
                            double* resolution);

/* Host copies of the tables (any may be NULL): mask (1 inside the scan),
 * sample_index (first of the neighbouring samples, 0xffffffff outside) and
 * weights [3][pixels] (x, depth, y weight; the y weight is 0 in 2D).
 * Synchronous. */
int supra_scanconv_tables(const supra_scanconv* sc, uint8_t* mask, uint32_t* sample_index, float* weights);

/* ScanConverter::getMask (ScanConverter.cu:430-433), the node's mask output:
 * writes the uint8 mask to device memory, asynchronously on `stream`. */
int supra_scanconv_mask(const supra_scanconv* sc, uint8_t* mask_device, void* stream);

/* ScanConverter::convert<InputType, OutputType> (ScanConverter.cu:436-488)
 * for num_frames consecutive frames: in [num_frames][num_samples][scanlines],
 * out [num_frames][pixels].  in_dtype and out_dtype each SUPRA_UINT8,
 * SUPRA_INT16 or SUPRA_FLOAT32. */
int supra_scanconv_run(const supra_scanconv* sc, const void* scanlines_in, int32_t in_dtype, uint32_t num_frames,
                       void* out, int32_t out_dtype, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SUPRA_SCANCONV_H */
