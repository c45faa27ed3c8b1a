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

/* Union of solid primitives (the batched-solids workload): voxel (x, y, z)
 * is inside iff its centre lies in one of the n_prims primitives, and the
 * outer 1-voxel shell is empty.  prims (host) holds VC_SOLID_WORDS doubles per
 * primitive:
 *   [0] kind: 0 ellipsoid, 1 capsule, 2 torus   [1..3] centre c
 *   [4..12] rotation R, row-major; primitive-frame coordinates
 *           q_j = ((x-c0) R[0][j] + (y-c1) R[1][j]) + (z-c2) R[2][j]
 *   [13..15] ellipsoid: semi-axes a, b, c -- inside iff
 *            ((q0/a)^2 + (q1/b)^2) + (q2/c)^2 <= 1
 *            capsule: radius r, half-length h -- inside iff
 *            ((q0 - clamp(q0, -h, h))^2 + q1^2) + q2^2 <= r^2
 *            torus: major R, minor r (axis q2) -- inside iff
 *            (sqrt(q0^2 + q1^2) - R)^2 + q2^2 <= r^2
 * Every operation is one IEEE double operation in the order written.
 * Waits for the result. */
#define VC_SOLID_WORDS 16
int vc_synth_solids(vc_ctx* ctx, uint8_t* d_occ, int nx, int ny, int nz,
                    const double* prims, int n_prims);

#ifdef __cplusplus
}
#endif
#endif
