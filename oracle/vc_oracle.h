/*
 * vc_oracle.h -- CPU restatement of the Voxel Cores volumetric hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product library (libvoxelcore.so) never
 * links, loads or calls it.
 *
 * Volume layout used everywhere: uint8 occupancy, C order [x][y][z]
 * (z fastest), linear index  i = (x*ny + y)*nz + z.  Non-zero = inside.
 * This is the reference's own traversal order (x outer, z inner,
 * src/surfacing.cpp:275-283), so "linear index order" == "Surfacer order".
 *
 * Each function names the reference file:line it restates, or says
 * "project-defined" where the reference has no counterpart (the EDT, the
 * per-voxel measure and the voxel subfield thinning).
 */
#ifndef VC_ORACLE_H
#define VC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- exact squared EDT + feature transform (project-defined) */
/* d2[p] = min over background q of |p-q|^2 (0 on background); ft[p] = linear
 * index of the lowest-index minimiser; -1 in both when the volume has no
 * background voxel.  Separable lower envelopes, passes z, y, x. */
/* threads of the restatement's OpenMP loops (1 = sequential; results do
 * not depend on it); returns the number in use */
int or_set_threads(int n);

int or_edt_ft(const uint8_t* occ, int nx, int ny, int nz, int32_t* d2,
              int32_t* ft);
/* O(N*B) brute force (tiny volumes only). */
int or_edt_brute(const uint8_t* occ, int nx, int ny, int nz, int32_t* d2,
                 int32_t* ft);
/* 1-D lower envelope: D[j] = min_i f[i] + (j-i)^2 over f[i] >= 0, arg = lowest
 * minimising i; D = -1, arg = -1 when every f is < 0. */
void or_envelope_1d(const int64_t* f, int n, int64_t* D, int32_t* arg);

/* ---------------- per-voxel medial measure (project-defined, R16) ------- */
/* lam2[p] = max over 6-neighbours q inside the object of |ft[p]-ft[q]|^2
 * (0 for background or when ft is undefined).  The voxel analogue of
 * MeasureForMA::lambdaForFace (include/measureforMA_imp.h:1-4) applied to the
 * two feature sites of a voxel face, max-aggregated like
 * VoroInfo::computeEdgesMeasure (src/voroinfo.cpp:1490-1538). */
void or_lambda2(const uint8_t* occ, const int32_t* ft, int nx, int ny, int nz,
                int32_t* lam2);
/* Integer thresholds for float levels: T_k = ceil(level_k^2) so that
 * lambda < level  <=>  lam2 < T_k exactly. */
int64_t or_level_threshold(float level);
/* lvl[p] = smallest k with lam2[p] < T[k] (inside voxels), 255 otherwise. */
void or_levels(const uint8_t* occ, const int32_t* lam2, int64_t n,
               const int64_t* T, int nlev, uint8_t* lvl);

/* ---------------- simple points, (26,6) adjacency ---------------------- */
/* nb26 bit b <-> neighbour j = (dz+1) + 3(dy+1) + 9(dx+1), b = j - (j > 13). */
int or_is_simple(uint32_t nb26);            /* Bertrand-Malandain T26/T6 */
int or_is_simple_topo(uint32_t nb26);       /* b0(X), b0(~X), chi preserved */
void or_simple_lut(uint32_t* lut_words);    /* 2^21 words, bit = simple */
/* all 2^26 codes: BFS definition vs the bit-parallel one (mismatches);
 * every topo_stride-th code also vs the 5x5x5 topological definition */
int64_t or_simple_exhaustive(int64_t* n_simple, int64_t topo_stride,
                             int64_t topo_offset, int64_t* topo_mismatch,
                             int64_t* topo_checked);

/* ---------------- measure-ordered subfield thinning (project-defined) --- */
typedef struct or_thin_stats {
    int64_t removed[64]; /* voxels removed per level */
    int64_t sweeps[64];  /* full 8-subfield sweeps per level (incl. last) */
    int64_t tests;       /* simple-point tests performed */
} or_thin_stats;
/* label: in = occupancy (0/1), out = 0 background, 1 survives every level,
 * 2+k removed at level k.  lvl from or_levels. */
int or_thin(uint8_t* label, const uint8_t* lvl, int nx, int ny, int nz,
            int nlev, or_thin_stats* st);

/* ---------------- topology of a voxel set ------------------------------ */
/* chi of the closed cubical complex of the inside voxels (V-E+F-C). */
int64_t or_euler_voxels(const uint8_t* occ, int nx, int ny, int nz);
/* # 26-components of inside voxels; # 6-components of the complement of the
 * inside voxels within the grid padded by one background shell. */
int64_t or_components26(const uint8_t* occ, int nx, int ny, int nz);
int64_t or_bg_components6(const uint8_t* occ, int nx, int ny, int nz);

/* ---------------- Surfacer restatement (reference-pinned, R3/R4) ------- */
typedef struct or_euler {
    int64_t V, E, F, T, C, euler;
    int64_t V_bnd, E_bnd, V_int, E_int, F_int;
} or_euler;
/* 256-entry surface-component count per 2x2x2 mask (src/surfacing.cpp:458-552
 * semantics, derived from first principles). */
void or_cube_components(int32_t* tbl256);
/* Surfacer::extractBoundaryVtsAndFaces (src/surfacing.cpp:49-197) +
 * computeEulerChar (:323-440) + compute_conn_cmpnt_num (:619-660).
 * sites: 3 floats per site; quads: 4 int32 per quad, both in Surfacer's
 * discovery order.  Returns 0, or -1 when a capacity is too small (counts are
 * still written).  euler may be NULL. */
int or_boundary(const uint8_t* occ, int nx, int ny, int nz, float* sites,
                int64_t site_cap, int64_t* n_sites, int32_t* quads,
                int64_t quad_cap, int64_t* n_quads, or_euler* euler);

/* ---------------- cell complex + simple-pair collapse (reference-pinned) */
/* cellcomplex::finalize (src/cellcomplex.cpp:364-491): edge ids = the given
 * edges first (exact-pair dedup, first occurrence wins), then the new face
 * sides (makeEdge-normalised) in face order.  fe_out receives each face's
 * side edge ids (same offsets as fv).  Returns the final edge count, or -1
 * when edge_cap is too small. */
int or_cc_finalize(int nv, int ne_in, const int32_t* edges_in, int nf,
                   const int64_t* foff, const int32_t* fv, int32_t* edges_out,
                   int edge_cap, int32_t* fe_out);
/* CellComplexThinning::prune (src/ccthin.cpp:201-424, :514-771): FIFO
 * collapse of face-edge pairs (face measure < f_t) and edge-vertex pairs
 * (edge measure < l_t).  Outputs removed flags and final edge ref counts.
 * Returns 0, or -1 on the reference's "logic err" exit. */
int or_cc_prune(int nv, int ne, const int32_t* edges, int nf,
                const int64_t* foff, const int32_t* fe, const float* em,
                const float* fm, float f_t, float l_t, uint8_t* rm_v,
                uint8_t* rm_e, uint8_t* rm_f, int32_t* ref_e);
/* util::simpleTriangulate (src/geomalgo.cpp:13-35) of one polygon; returns
 * the number of triangles written (n - 2). */
int or_simple_triangulate(const int32_t* poly, int n, int32_t* tris);

#ifdef __cplusplus
}
#endif
#endif
