/*
 * supra_synth.h — seeded synthetic RF for tests and benchmarks.
 *
 * Produces USRawData-shaped int16 (or float) RF [tx][channel][sample] from
 * point scatterers: every (transmit, channel) line is the sum over scatterers
 * of a Gaussian-modulated pulse delayed by
 *     t0(tx) + (depth of the scatterer along the transmit axis
 *               + distance scatterer -> element) / c,
 * weighted by the scatterer amplitude and a Gaussian transmit-beam profile,
 * plus white Gaussian noise at a given SNR, then scaled to a given peak.
 * No reference counterpart on the beamforming path (the reference's own
 * simulated sources are UltrasoundInterfaceSimulated / NoiseNode, which are
 * not part of this library).  Deterministic: counter-based random numbers,
 * fixed-point accumulation of the echoes.
 */
#ifndef SUPRA_SYNTH_H
#define SUPRA_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct supra_synth_params {
    uint32_t num_tx, num_channels, num_samples, num_elements;
    double   sampling_frequency_hz;
    double   speed_of_sound_m_per_s;
    double   center_frequency_hz;
    double   fractional_bandwidth;   /* -6 dB, e.g. 0.6 */
    double   snr_db;                 /* noise level relative to the signal RMS */
    double   peak;                   /* output scaled so max |value| == peak */
    double   beam_sigma_mm;          /* Gaussian transmit-beam width */
    uint64_t noise_seed;
    uint32_t num_scatterers;
    const double* scatterers;        /* [num_scatterers][4]: x, depth, z [mm], amplitude */
    const double* tx_origin;         /* [num_tx][3] mm (x, depth, z) */
    const double* tx_direction;      /* [num_tx][3] unit (x, depth, z) */
    const double* tx_t0;             /* [num_tx] s */
    const uint32_t* tx_first_element;/* [num_tx] inclusive receive aperture (linear index) */
    const uint32_t* tx_last_element; /* [num_tx] */
    const double* element_pos;       /* [num_elements][3] mm (x, depth=0, z) */
} supra_synth_params;

/* Uniform [0,1) from (seed, index): the generator both synthesizers use for
 * scatterer tables (exported so that host code draws the same numbers). */
double supra_synth_uniform(uint64_t seed, uint64_t index);

/* Host version (OpenMP over lines).  out: host [tx][ch][s], int16 if
 * out_dtype == SUPRA_INT16 (1) else float.  Returns 0 or a negative code. */
int supra_synth_rf_host(const supra_synth_params* p, void* out, int32_t out_dtype);

/* Device version: same model on the current HIP device (out: device
 * buffer; scratch of num_tx*num_channels*num_samples int32 is allocated and
 * freed internally).  Synchronises the stream before returning. */
int supra_synth_rf_device(const supra_synth_params* p, void* out, int32_t out_dtype, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SUPRA_SYNTH_H */
