/*
 * vc_cc.h -- a finalized 2-complex resident on the device, and the
 * reference's measure-thresholded collapse, remaining complex and topology
 * checks on it (libvoxelcore.so).
 *
 * A complex is created from host arrays once (vertex count, edges, faces as
 * vertex polygons with int64 offsets), finalized on the device and kept
 * there with its adjacency; a collapse state (vc_ccthin) over it runs any
 * number of prune calls, each from the full complex, as the threshold loop
 * of exportInsideVoroMesh does, without uploading the complex again.
 *
 * Reference entry points replaced (reference file:line):
 *   vc_cc_create        cellcomplex(vts, edges, faces) + finalize
 *                       (src/cellcomplex.cpp:13-26, :364-491)
 *   vc_cc_get           the finalized edge list and per-face side edge ids
 *                       (include/cellcomplex.h:132-148)
 *   vc_cc_topology      cellcomplex::eulerChar (src/cellcomplex.cpp:493-512,
 *                       util::findNumConnComponents src/geomalgo.cpp:67-116)
 *                       and cellcomplex::compute_conn_cmpnts (:74-165)
 *   vc_cc_subset_components  cellcomplex::compute_conn_cmpnts(subset, 1)
 *                       (:166-219), used by VoroInfo::computeEulerChar
 *                       (src/voroinfo.cpp:1386-1430)
 *   vc_ccthin_create    CellComplexThinning::setup + preprocess
 *                       (src/ccthin.cpp:27, :59-115)
 *   vc_ccthin_assign    CellComplexThinning::assignElementValues
 *                       (src/ccthin.cpp:178-189; measures copied)
 *   vc_ccthin_prune     CellComplexThinning::prune (src/ccthin.cpp:201-424,
 *                       :514-771), the FIFO order reproduced exactly
 *   vc_ccthin_removed   getEdgeRemovedTags / getFaceRemovedTags and the
 *                       vertex tags (src/ccthin.cpp:113-125)
 *   vc_ccthin_remaining CellComplexThinning::remainingCC
 *                       (src/ccthin.cpp:426-512)
 *   vc_ccthin_remained  getRemainedVts / getRemainedEdges / getRemainedFaces
 *                       (src/ccthin.cpp:127-176)
 *   vc_ccthin_remaining_euler  util::computeEulerChar(vts, edges, tris)
 *                       (src/geomalgo.cpp:174-214) of the remaining complex
 *   vc_cc_closed_components  voxelvoro::nClosedComponents
 *                       (src/highlevelalgo.cpp:993-1020)
 *   vc_euler_char       util::computeEulerChar(vts, edges, tris) of arrays
 */
#ifndef VC_CC_H
#define VC_CC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vc_ctx vc_ctx;
typedef struct vc_cc vc_cc;
typedef struct vc_ccthin vc_ccthin;

/* eulerchar (include/commondefs.h:60-85) of a complex, plus the number of
 * face components through shared edges (-1 where not computed) */
typedef struct vc_cc_topo {
    int64_t V, E, F, T, C, euler;
    int64_t face_components;
} vc_cc_topo;

/* cellcomplex ctor + finalize on the device: edge ids are the given edges
 * first (exact-pair dedup, first occurrence wins), then the new face sides,
 * makeEdge-normalised, in face order (vc_cc_finalize's numbering).  Host
 * arrays, copied; VC_ERR_INVALID for ids out of range.  Waits. */
int vc_cc_create(vc_ctx* ctx, int32_t nv, int32_t ne_in,
                 const int32_t* edges_in, int32_t nf, const int64_t* face_off,
                 const int32_t* face_vts, vc_cc** out);
int vc_cc_release(vc_ctx* ctx, vc_cc* cc);
/* counts of the finalized complex; n_sides = total face sides */
int vc_cc_sizes(const vc_cc* cc, int32_t* nv, int32_t* ne, int32_t* nf,
                int64_t* n_sides);
/* finalized edges (2 * ne ints) and each face's side edge ids (n_sides
 * ints, face order, each face's sides in polygon order); either may be NULL;
 * waits */
int vc_cc_get(vc_ctx* ctx, vc_cc* cc, int32_t* edges, int32_t* face_edges);
/* eulerChar: V, E, F, T = 0, euler = V - E + F, C = vertex components through
 * the edges (every vertex counts, isolated ones as their own component);
 * face_components = compute_conn_cmpnts(): faces joined through shared
 * edges.  Union-find on the device; waits. */
int vc_cc_topology(vc_ctx* ctx, vc_cc* cc, vc_cc_topo* out);
/* compute_conn_cmpnts(subset, 1) (src/cellcomplex.cpp:166-219), the C of
 * VoroInfo::computeEulerChar after pruning (src/voroinfo.cpp:1386-1430):
 * components of the vertices of the faces with face_in[f] != 0 (nf bytes,
 * host), joined by those faces' polygon sides; other vertices take no part.
 * Waits. */
int vc_cc_subset_components(vc_ctx* ctx, vc_cc* cc, const uint8_t* face_in,
                            int64_t* n_comp);

/* collapse state over cc (cc must outlive it); measures start at 0 */
int vc_ccthin_create(vc_ctx* ctx, vc_cc* cc, vc_ccthin** out);
int vc_ccthin_release(vc_ctx* ctx, vc_ccthin* t);
/* edge and face measures (ne and nf floats, host, copied; waits) */
int vc_ccthin_assign(vc_ctx* ctx, vc_ccthin* t, const float* edge_measure,
                     const float* face_measure);
/* prune(f_t, l_t): from the full complex, queue every face-edge pair with
 * face measure < f_t (edge order), then every edge-vertex pair with edge
 * measure < l_t (vertex order), pop FIFO (one queue generation at a time,
 * as vc_cc_prune_device).  VC_ERR_STATE where the reference prints "logic
 * err" and exits; VC_ERR_INVALID for a complex with an edge whose ends are
 * equal or a face listing a side twice (vc_cc_prune handles those on the
 * host).  The results stay on the device.  generations may be NULL.
 * Waits. */
int vc_ccthin_prune(vc_ctx* ctx, vc_ccthin* t, float f_t, float l_t,
                    int64_t* generations);
/* removed flags (0/1) of vertices, edges, faces and the final edge
 * reference counts (faces left on each edge); any may be NULL; waits */
int vc_ccthin_removed(vc_ctx* ctx, vc_ccthin* t, uint8_t* removed_v,
                      uint8_t* removed_e, uint8_t* removed_f, int32_t* ref_e);
/* remainingCC: vts_idx = kept vertex ids ascending; edges = the kept
 * isolated edges (not removed, no face left on them) in edge order; tris =
 * the zig-zag triangles of each kept face (util::simpleTriangulate) in face
 * order, from_fi = their face; edge ends and corners renumbered by their
 * rank in vts_idx (util::compactify; an id that is not kept becomes 0, as
 * the reference's map lookup gives).  With every output NULL only the counts
 * are written; VC_ERR_CAPACITY (counts written) when a capacity is short.
 * Computed on the device; waits. */
int vc_ccthin_remaining(vc_ctx* ctx, vc_ccthin* t, int32_t* vts_idx,
                        int64_t vts_cap, int64_t* n_vts, int32_t* edges,
                        int64_t edge_cap, int64_t* n_edges, int32_t* tris,
                        int32_t* from_fi, int64_t tri_cap, int64_t* n_tris);
/* getRemainedVts / getRemainedEdges / getRemainedFaces: ids of the kept
 * vertices, the kept isolated edges and the kept faces, ascending.  With
 * every output NULL only the counts are written; VC_ERR_CAPACITY (counts
 * written) when a capacity is short.  Waits. */
int vc_ccthin_remained(vc_ctx* ctx, vc_ccthin* t, int32_t* v_ids,
                       int64_t v_cap, int64_t* n_v, int32_t* e_ids,
                       int64_t e_cap, int64_t* n_e, int32_t* f_ids,
                       int64_t f_cap, int64_t* n_f);
/* util::computeEulerChar(vts, edges, tris) of the remaining complex, on the
 * device without copying it out: V = kept vertices, E = distinct pairs among
 * the kept edges (as given) and the triangle sides (normalised), F =
 * triangles, C = vertex components through those pairs; face_components =
 * -1.  Waits. */
int vc_ccthin_remaining_euler(vc_ctx* ctx, vc_ccthin* t, vc_cc_topo* out);
/* nClosedComponents: prune with every measure 0 and thresholds 1, then the
 * components of the remaining triangles joined through shared (normalised)
 * sides; 0 when no face remains.  Waits. */
int vc_cc_closed_components(vc_ctx* ctx, vc_cc* cc, int64_t* n_closed);
/* util::computeEulerChar(vts, edges, tris) of host arrays (nv vertices, ne
 * edges, nt triangles; ids must be < nv), on the device; waits */
int vc_euler_char(vc_ctx* ctx, int32_t nv, const int32_t* edges, int64_t ne,
                  const int32_t* tris, int64_t nt, vc_cc_topo* out);

#ifdef __cplusplus
}
#endif
#endif /* VC_CC_H */
