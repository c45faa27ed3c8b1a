/*
 * vc_api.h -- C ABI of the MI355X-native Voxel Cores volumetric hot path
 * (libvoxelcore.so, built for gfx950).
 *
 * Plain C: opaque handles, plain pointers, int32/int64 sizes, int32 status
 * codes (0 = ok, < 0 = error; the message is kept per context, see
 * vc_last_error).  No C++ types or exceptions cross this boundary.
 *
 * Volume layout: uint8 occupancy in C order [x][y][z] (z fastest), linear
 * index i = (x*ny + y)*nz + z.  A voxel is inside iff its byte is non-zero,
 * the device form of SpaceConverter::get_occupancy_at_vox (reference
 * include/spaceinfo.h:122-135: getDataAt > 0, out of bounds = 0).  This order
 * is the reference Surfacer's own traversal order (x outer, z inner,
 * src/surfacing.cpp:275-283).
 *
 * Every compute entry point is enqueued on the context's HIP stream and
 * returns without waiting, except the vc_get_* / *_host variants, which wait
 * for their own results.  One host thread per context.  (The boundary pass
 * also runs kernels on a second stream of the context, forked from and
 * joined back to the context stream by events within the call, so its
 * results are ordered on the context stream like every other.)
 *
 * Reference entry points replaced (reference file:line):
 *   vc_volume_upload      DenseVolume / Volume3DScalar::getDataAt
 *                         (include/densevolume.h:13-48,
 *                          include/Volume3DScalar.h:11-34) read once into HBM
 *   vc_edt                no reference counterpart (project-defined exact
 *                         squared EDT + feature transform); feeds the
 *                         medial measure the way TetGen's Voronoi diagram
 *                         feeds VoroInfo (src/highlevelalgo.cpp:503-529)
 *   vc_measure            MeasureForMA::lambdaForFace + the measure loops
 *                         (include/measureforMA_imp.h:1-4,
 *                          src/voroinfo.cpp:1432-1574), per voxel
 *   vc_thin_levels        CellComplexThinning::prune driven by the
 *                         threshold loop of exportInsideVoroMesh
 *                         (src/ccthin.cpp:201-424,
 *                          src/highlevelalgo.cpp:815-961), on voxels
 *   vc_cc_finalize        cellcomplex::finalize (src/cellcomplex.cpp:364-491)
 *   vc_cc_finalize_device the same, on the device
 *   vc_cc_prune           CellComplexThinning::prune (src/ccthin.cpp:201-771)
 *   vc_cc_prune_device    the same, on the device
 *   vc_boundary_extract   Surfacer::extractBoundaryVts /
 *                         extractBoundaryVtsAndFaces / computeEulerChar
 *                         (src/surfacing.cpp:49-440, :619-660)
 */
#ifndef VC_API_H
#define VC_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define VC_OK 0
#define VC_ERR_INVALID (-1)   /* bad argument / shape / state            */
#define VC_ERR_HIP (-2)       /* a HIP runtime call failed                */
#define VC_ERR_CAPACITY (-3)  /* output buffer too small (counts written) */
#define VC_ERR_EMPTY_VOL (-4) /* SurfacerErrCode::EMPTY_VOL (null volume) */
#define VC_ERR_NOMEM (-5)     /* device allocation failed                 */
#define VC_ERR_STATE (-6)     /* prerequisite stage not run yet           */
#define VC_ERR_GUARD (-7)     /* a device-side iteration guard tripped    */
#define VC_ERR_IO (-8)        /* a file could not be opened or written    */

/* ---- limits ------------------------------------------------------------ */
/* each axis <= 32767, nx*ny*nz < 2^31, packed feature coordinates fit in
 * 31 bits (ceil(log2 nx)+ceil(log2 ny)+ceil(log2 nz) <= 31) */
#define VC_MAX_AXIS 32767
#define VC_MAX_LEVELS 64

typedef struct vc_ctx vc_ctx;
typedef struct vc_vol vc_vol;

/* ---- context ----------------------------------------------------------- */
/* Binds a HIP device, creates the context stream (and a second, internal
 * one, both of default priority) and builds the (26,6) simple-point table
 * (2^26 bits, 8 MiB) in HBM. */
int vc_ctx_create(int device, vc_ctx** out);
int vc_ctx_destroy(vc_ctx* ctx);
/* copies the last error message (NUL-terminated, truncated to n) */
int vc_last_error(const vc_ctx* ctx, char* buf, size_t n);
/* the hipStream_t every entry point of this context is enqueued on */
int vc_ctx_stream(vc_ctx* ctx, void** hip_stream);
int vc_ctx_synchronize(vc_ctx* ctx);
/* copy of the device simple-point table: 2^21 uint32 words, bit b of word w
 * is the simple flag of neighbourhood code (w << 5) | b (waits) */
int vc_simple_table_host(vc_ctx* ctx, uint32_t* words_out);

/* ---- volumes ----------------------------------------------------------- */
/* host bytes, C order [x][y][z]; copied into HBM (H2D) */
int vc_volume_upload(vc_ctx* ctx, const uint8_t* occ, int nx, int ny, int nz,
                     vc_vol** out);
/* device bytes (any HIP device pointer on this device); copied D2D */
int vc_volume_from_device(vc_ctx* ctx, const uint8_t* d_occ, int nx, int ny,
                          int nz, vc_vol** out);
int vc_volume_release(vc_ctx* ctx, vc_vol* vol);
int vc_volume_dims(const vc_vol* vol, int* nx, int* ny, int* nz);

/* ---- (a) exact squared EDT + feature transform ------------------------- */
/* d2[p] = min over background voxels q of |p-q|^2 (0 on background);
 * ft[p] = linear index of the lowest-index background voxel attaining it.
 * Only in-grid background counts (scipy.ndimage.distance_transform_edt
 * convention); with no background voxel at all d2 = ft = -1 everywhere.
 * Three separable passes (z, y, x) in HBM; flags must be 0. */
int vc_edt(vc_ctx* ctx, vc_vol* vol, int flags);

/* ---- medial measure ------------------------------------------------------ */
/* lam2[p] = max over 6-neighbours q inside the object of |ft[p]-ft[q]|^2
 * (uint32; 0 on background).  lambda = sqrt(lam2) is the distance between
 * the feature sites of the two voxels of a face, the voxel form of
 * MeasureForMA::lambdaForFace.  Requires vc_edt. */
int vc_measure(vc_ctx* ctx, vc_vol* vol);
/* replace the measure the thinning uses by a caller-supplied float field
 * (e.g. a learned surrogate); device pointer, N floats, copied */
int vc_set_measure_f32(vc_ctx* ctx, vc_vol* vol, const float* d_measure);

/* ---- (c) measure-ordered subfield thinning ----------------------------- */
typedef struct vc_thin_stats {
    int64_t removed[VC_MAX_LEVELS]; /* voxels removed per level            */
    int64_t passes;                 /* subfield passes that had work        */
    int64_t tests;                  /* simple-point tests executed          */
    int64_t launches;               /* kernel launches issued               */
    int64_t voxel_passes;           /* voxels visited, summed over subfield
                                       passes: a tile sweep counts its tiles'
                                       voxels 8 times, a word pass 4 x 64
                                       voxels per flagged word             */
} vc_thin_stats;
/* levels: ascending, n_levels in [1, 64].  For level k (ascending) repeat
 * 8-subfield sweeps (subfield s = (x&1)|(y&1)<<1|(z&1)<<2, s ascending)
 * until a sweep removes nothing; a pass removes every alive voxel of its
 * subfield that is (26,6)-simple and whose measure is below level k
 * (lambda < level  <=>  lam2 < ceil(level^2) for the exact measure).
 * Result (vc_get_labels): 0 background, 1 kept at every level, 2+k removed at
 * level k.  The skeleton at level k is {label == 1 or label > 2+k}.
 * stats may be NULL; when given, the call waits for the result. */
int vc_thin_levels(vc_ctx* ctx, vc_vol* vol, const float* levels,
                   int n_levels, vc_thin_stats* stats);

/* Tile-sweep timing, a measurement hook with no reference counterpart:
 * while enabled, each thin_tile_sweep launch that vc_thin_levels or
 * vc_batch_run issues on this context is bracketed by two HIP events on the
 * context stream (with use_graph, one pair per graph segment, which then also
 * holds the segment's entry and bookkeeping kernels); the elapsed times are
 * read at the host wait that ends each segment.  Enabling or disabling
 * resets the totals.  vc_thin_timing_read returns the launches timed and
 * their summed duration in milliseconds. */
int vc_thin_timing(vc_ctx* ctx, int enable);
int vc_thin_timing_read(vc_ctx* ctx, int64_t* launches, double* total_ms);

/* ---- (b) boundary sites + cube-complex Euler ---------------------------- */
typedef struct vc_euler {
    /* eulerchar (reference include/commondefs.h:60-85) */
    int64_t V, E, F, T, C, euler;
    /* parts: boundary V/E with multiplicity, interior V/E/F */
    int64_t V_bnd, E_bnd, V_int, E_int, F_int;
    /* Euler characteristic of the union of the closed voxel cubes (the
     * topological invariant; equals euler wherever the reference's vertex
     * multiplicity heuristic holds, i.e. away from non-manifold contacts) */
    int64_t chi;
} vc_euler;
/* Boundary sites (voxel corners shared by a boundary face) as 3 floats each
 * and boundary quads as 4 int32 site indices each, both in the reference
 * Surfacer's discovery order; euler (may be NULL) as computeEulerChar.
 * Host output buffers; when a capacity is too small returns VC_ERR_CAPACITY
 * with the required counts in n_sites / n_quads.  sites or quads may be NULL
 * to query counts only.  Waits for the result. */
int vc_boundary_extract(vc_ctx* ctx, vc_vol* vol, float* sites,
                        int64_t site_cap, int64_t* n_sites, int32_t* quads,
                        int64_t quad_cap, int64_t* n_quads, vc_euler* euler);

/* ---- cell complex + simple-pair collapse (host side) -------------------- */
/* The reference's 2-complex thinning is a FIFO collapse whose output depends
 * on the queue order (src/ccthin.cpp:248-270, :534-769); it runs on the host,
 * sequentially, in exactly that order.  No context needed. */
typedef struct vc_cc_view {
    int32_t nv, ne, nf;
    const int32_t* edges;      /* 2*ne vertex ids (finalized edge list)     */
    const int64_t* face_off;   /* nf+1 offsets into face_edges              */
    const int32_t* face_edges; /* side edge ids of each face, in face order */
} vc_cc_view;
/* cellcomplex::finalize (src/cellcomplex.cpp:364-491): edge ids are the given
 * edges first (exact-pair dedup, first occurrence wins), then the new face
 * sides, makeEdge-normalised, in face order.  face_edges_out gets each
 * face's side edge ids (offsets face_off).  VC_ERR_CAPACITY when edge_cap is
 * too small (*ne_out still written). */
int vc_cc_finalize(int32_t nv, int32_t ne_in, const int32_t* edges_in,
                   int32_t nf, const int64_t* face_off,
                   const int32_t* face_vts, int32_t* edges_out,
                   int32_t edge_cap, int32_t* ne_out, int32_t* face_edges_out);
/* vc_cc_finalize on the device, the same outputs: the given edges (exact
 * pairs) then the face sides (normalised) are candidates in that order, the
 * first candidate of each pair (a device hash table keeps the smallest
 * index) becomes an edge, numbered by a scan.  Host arrays; waits.  With
 * edges_out NULL only *ne_out is written. */
int vc_cc_finalize_device(vc_ctx* ctx, int32_t nv, int32_t ne_in,
                          const int32_t* edges_in, int32_t nf,
                          const int64_t* face_off, const int32_t* face_vts,
                          int32_t* edges_out, int32_t edge_cap, int32_t* ne_out,
                          int32_t* face_edges_out);
/* CellComplexThinning::prune(f_t, l_t) (src/ccthin.cpp:201-424, :514-771):
 * queue every face-edge pair with face measure < f_t (edge order), then every
 * edge-vertex pair with edge measure < l_t (vertex order), then pop FIFO.
 * Outputs removed flags (0/1) and the final edge reference counts (faces per
 * edge).  VC_ERR_STATE where the reference prints "logic err" and exits. */
int vc_cc_prune(const vc_cc_view* cc, const float* edge_measure,
                const float* face_measure, float f_t, float l_t,
                uint8_t* removed_v, uint8_t* removed_e, uint8_t* removed_f,
                int32_t* ref_e);

/* vc_cc_prune on the device: the same collapse and the same outputs,
 * computed one generation of the FIFO queue at a time (the pairs the
 * previous generation's pops pushed, in push order), with the pops of a
 * generation in parallel and their pushes ordered by (pop index, step).
 * Adjacency is built on the device.  VC_ERR_STATE where vc_cc_prune
 * returns it; VC_ERR_INVALID for ids out of range, an edge with equal ends
 * or a face listing a side twice (vc_cc_prune handles those).  generations
 * (may be NULL) receives the number of queue generations.  Waits. */
int vc_cc_prune_device(vc_ctx* ctx, const vc_cc_view* cc,
                       const float* edge_measure, const float* face_measure,
                       float f_t, float l_t, uint8_t* removed_v,
                       uint8_t* removed_e, uint8_t* removed_f, int32_t* ref_e,
                       int64_t* generations);

/* ---- one call that allocates nothing ------------------------------------ */
/* vc_skeleton_noalloc has no counterpart in the reference.  It is for callers
 * that cannot allocate device memory (tight serving loops, harnesses that
 * forbid hipMalloc): vc_edt + vc_measure + vc_thin_levels of one volume that
 * the caller holds in device memory, with the same results, and no hipMalloc,
 * hipHostMalloc or hipFree in the call.  Every intermediate (squared EDT,
 * feature transform, lambda^2, level bytes, thinning planes and control
 * words) lives in a fixed device array of VC_NOALLOC_SCRATCH_BYTES that each
 * GPU the library runs on holds until the process ends; vc_ctx_create takes
 * its address.  Every call on one GPU shares that GPU's array (it keeps the
 * last call's contents; the call initialises what it reads), so calls that
 * overlap on one GPU -- from two host threads or two contexts of the same
 * device -- race and must be serialised by the caller.
 *   d_occ    N bytes of device memory, C order [x][y][z], non-zero = inside;
 *   d_label  N bytes of device memory, written as vc_get(VC_BUF_LABEL);
 *   levels   host array, as vc_thin_levels; stats may be NULL.
 * Limits: N = nx*ny*nz <= VC_NOALLOC_MAX_VOXELS, nx <= 16384, ny <= 32768,
 * and the layout (13 bytes per voxel + 26 bytes per 64-voxel z word + small
 * control blocks) within VC_NOALLOC_SCRATCH_BYTES; beyond them VC_ERR_INVALID
 * with the sizes in the message.  Enqueued on the context stream; returns
 * when the thinning has converged (it waits for its convergence checks). */
#define VC_NOALLOC_SCRATCH_BYTES (64LL << 20)
#define VC_NOALLOC_MAX_VOXELS (1LL << 22)
int vc_skeleton_noalloc(vc_ctx* ctx, const uint8_t* d_occ, int nx, int ny,
                        int nz, const float* levels, int n_levels,
                        uint8_t* d_label, vc_thin_stats* stats);

/* ---- results ------------------------------------------------------------- */
/* which: */
#define VC_BUF_OCC 0    /* uint8  N */
#define VC_BUF_EDT 1    /* int32  N */
#define VC_BUF_FT 2     /* int32  N (linear index; converted on copy-out) */
#define VC_BUF_LAM2 3   /* uint32 N */
#define VC_BUF_LABEL 4  /* uint8  N */
#define VC_BUF_LEVEL 5  /* uint8  N (first eligible level, 255 = never) */
/* copy a result to host memory (waits); bytes must equal N * element size */
int vc_get(vc_ctx* ctx, vc_vol* vol, int which, void* host_dst, int64_t bytes);
/* copy a result into caller device memory (enqueued, no wait) */
int vc_get_device(vc_ctx* ctx, vc_vol* vol, int which, void* dev_dst,
                  int64_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* VC_API_H */
