/*
 * vc_synth.h -- seeded synthetic inputs generated on the device
 * (libvoxelcore.so).  Benchmark plumbing: no reference counterpart; the
 * reference reads its volumes from .mrc/.sof/.sog files
 * (src/importers.cpp:19-77).
 */
#ifndef VC_SYNTH_H
#define VC_SYNTH_H

#include <stdint.h>

#include "vc_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Fills d_occ (device, nx*ny*nz bytes, C order [x][y][z]) with the
 * blurred-noise volume: N(0,1) noise (splitmix64 + Box-Muller, counter
 * based), separable Gaussian blur (2*round(4*sigma)+1 taps, periodic), inside
 * iff value > the exact ceil((1-fill)*N)-th smallest value; shell != 0 clears
 * the outer 1-voxel shell.  threshold_out (may be NULL) receives the
 * threshold value.  Waits for the result. */
int vc_synth_blurred_noise(vc_ctx* ctx, uint8_t* d_occ, int nx, int ny,
                           int nz, float sigma, float fill, uint64_t seed,
                           int shell, float* threshold_out);

#ifdef __cplusplus
}
#endif
#endif
