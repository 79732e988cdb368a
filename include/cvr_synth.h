/*
 * cvr_synth.h — seeded synthetic volumes for tests and benchmarks (host code).
 *
 * The reference renders datasets read from disk (datamanager.cpp:69-101); these
 * generators stand in for them with reproducible inputs of the same form: a
 * uint8/uint16 grid, x fastest.  Random numbers come from SplitMix64
 * (Steele, Lea, Flood, "Fast splittable pseudorandom number generators", 2014).
 */
#ifndef CVR_SYNTH_H
#define CVR_SYNTH_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Sum of `n_blobs` isotropic Gaussians in normalised coordinates u = (i+0.5)/n:
 * centres U[0.2,0.8]^3, sigma U[0.05,0.15], amplitude U[0.3,1.0] (drawn in that
 * order per blob), plus `noise_weight` times value noise (see below; 0 disables),
 * normalised to [0, 255] (dtype 0) or [0, 65535] (dtype 1) and rounded. */
int cvr_synth_blobs(int nx, int ny, int nz, int dtype, int n_blobs, double noise_weight, int noise_octaves,
                    uint64_t seed, void* out);

/* `octaves` of value noise (lattice 8^3 at octave 0, doubling, gain 0.5,
 * smoothstep-weighted trilinear), normalised to [0,1]; if `n_bands` > 0 the field
 * is cut at its 1/n_bands quantiles into bands whose values are `band_values`
 * (fractions of the dtype maximum), else it is scaled to the dtype range. */
int cvr_synth_noise(int nx, int ny, int nz, int dtype, int octaves, uint64_t seed, int n_bands,
                    const double* band_values, void* out);

#ifdef __cplusplus
}
#endif
#endif
