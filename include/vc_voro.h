/*
 * vc_voro.h -- the Voronoi-side steps of the pipeline (libvoxelcore.so):
 * inside tagging, lambda measures on a 2-complex, inside extraction with
 * measures and close-vertex merging, on plain arrays.
 *
 * A complex is given as vertex coordinates (3 floats each), an edge list
 * (2 int32 each, the finalized cellcomplex numbering) and faces as vertex
 * polygons (face_off: nf+1 int64 offsets into face_vts).  Adjacency, where a
 * function needs it, is CSR in the cellcomplex order: edges of a vertex in
 * edge-id order, faces of an edge in face-id order.
 *
 * Reference entry points replaced (reference file:line):
 *   vc_voro_tag_points   VoroInfo::tagVert / tagVtsUsingUniformVol
 *                        (src/voroinfo.cpp:447-494; SpaceConverter
 *                        fromModelToVox / voxTaggedAsInside,
 *                        include/spaceinfo.h:53-58, :93-105)
 *   vc_voro_measures     VoroInfo::computeFacesMeasure / computeEdgesMeasure
 *                        / computeVertexMeasure with MeasureForMA::LAMBDA
 *                        (src/voroinfo.cpp:1432-1574,
 *                        include/measureforMA_imp.h:1-4)
 *   vc_voro_extract_inside  VoroInfo::extractInsideWithMeasure
 *                        (src/voroinfo.cpp:624-729); also on the device
 *                        (vc_voro_extract_inside_device)
 *   vc_voro_merge_close  VoroInfo::mergeCloseVts (src/voroinfo.cpp:1064-1384);
 *                        vc_voro_merge_close_device when nothing merges
 *   vc_voro_radii        the radius step of VoroInfo::computeInfoRelatedtoSites
 *                        and of the TetGen loader (src/voroinfo.cpp:290-325,
 *                        :404-441)
 *   vc_voro_part_inside  VoroInfo::getPartInside (src/voroinfo.cpp:573-620)
 *   vc_voro_from_ft      computeVD's TetGen Voronoi diagram of the boundary
 *                        sites (src/highlevelalgo.cpp:487-529), as the
 *                        discrete diagram of the feature transform
 */
#ifndef VC_VORO_H
#define VC_VORO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vc_ctx vc_ctx;
typedef struct vc_vol vc_vol;

/* valid[i] = 1 iff the voxel round(pts[i]) (each coordinate rounded half
 * away from zero) is in the volume and inside.  Host arrays; runs on the
 * device; waits. */
int vc_voro_tag_points(vc_ctx* ctx, vc_vol* vol, const float* pts, int64_t n,
                       uint8_t* valid);

/* Lambda measures (host arrays, computed on the device, waits):
 *   lam_f[f] = |site[face_sites[2f]] - site[face_sites[2f+1]]| (float:
 *              ((d0^2 + d1^2) + d2^2), correctly rounded sqrt), every face;
 *   lam_e[e] = max over faces f of e with face_valid[f] of lam_f (0 if none);
 *   lam_v[v] = max over edges e of v, over faces f of e with face_valid[f],
 *              of lam_f (0 if none).
 * Any of lam_v / lam_e / lam_f may be NULL. */
int vc_voro_measures(vc_ctx* ctx, int32_t nv, int32_t ne, int32_t nf,
                     const float* sites, int32_t nsites,
                     const int32_t* face_sites, const uint8_t* face_valid,
                     const int64_t* ve_off, const int32_t* ve,
                     const int64_t* ef_off, const int32_t* ef, float* lam_v,
                     float* lam_e, float* lam_f);

/* Inside extraction (host): with valid vertices V (ascending), valid edges
 * and valid faces (a face is valid iff all its vertices are; an edge iff both
 * ends are), outputs
 *   vts_idx  valid vertex ids (n_vts), in ascending order;
 *   edges    the valid edges (as stored), then the sides of the triangles of
 *            each valid face (zig-zag triangulation, util::simpleTriangulate)
 *            not seen before (exact pair, makeEdge-normalised sides), in face
 *            then triangle then side order, renumbered into vts_idx order;
 *   tris     the triangles of each valid face, renumbered;
 *   from_fi  the face each triangle comes from;
 *   em       valid-edge measures (lam_e of each valid edge), then the face's
 *            measure for every appended edge;
 *   fm       the face's measure for every triangle.
 * Capacities: vts_cap, edge_cap, tri_cap; VC_ERR_CAPACITY (-3) with the
 * required counts written when one is too small; outputs may be NULL for a
 * size query. */
int vc_voro_extract_inside(int32_t nv, int32_t ne, int32_t nf,
                           const int32_t* edges, const int64_t* face_off,
                           const int32_t* face_vts, const uint8_t* v_valid,
                           const float* lam_e, const float* lam_f,
                           int32_t* vts_idx, int64_t vts_cap, int64_t* n_vts,
                           int32_t* edges_out, float* em, int64_t edge_cap,
                           int64_t* n_edges, int32_t* tris, int32_t* from_fi,
                           float* fm, int64_t tri_cap, int64_t* n_tris);

/* vc_voro_extract_inside on the device: the same outputs in the same order
 * (valid vertices by compaction, valid edges by compaction, triangles at
 * scanned offsets, each appended side kept iff it is the first of its exact
 * pair in face / triangle / side order and no valid edge is stored as that
 * pair: a hash table keeps the smallest candidate index per pair).  Host
 * arrays; waits. */
int vc_voro_extract_inside_device(
    vc_ctx* ctx, int32_t nv, int32_t ne, int32_t nf, const int32_t* edges,
    const int64_t* face_off, const int32_t* face_vts, const uint8_t* v_valid,
    const float* lam_e, const float* lam_f, int32_t* vts_idx, int64_t vts_cap,
    int64_t* n_vts, int32_t* edges_out, float* em, int64_t edge_cap,
    int64_t* n_edges, int32_t* tris, int32_t* from_fi, float* fm,
    int64_t tri_cap, int64_t* n_tris);

/* Close-vertex merge (host): for each vertex vi in order (skipping visited,
 * and, with only_inside, invalid ones) start a new vertex and absorb by BFS
 * over the vertex-vertex adjacency (ve_off / ve in edge order) every
 * unvisited neighbour within eps of vi (squared float distance <= eps^2).
 * Outputs:
 *   old_to_new  nv entries (-1 for vertices never reached);
 *   new_vts     the seed of every new vertex (n_new_vts entries, <= nv);
 *   faces       for every kept face (>= 3 distinct new vertices, first
 *               occurrence order), its new polygon (new_face_off /
 *               new_face_vts, caps nf+1 / total face size) and
 *               remove_face[f] = 0; degenerate faces get remove_face = 1;
 *   edges       the set of new edges: kept edges (u != v, only_inside:
 *               valid) and faces degenerated to two vertices, makeEdge-
 *               normalised, deduplicated, sorted ascending (the reference
 *               appends them in std::unordered_set iteration order, which is
 *               implementation-defined; ascending order is this build's
 *               canonical choice), cap edge_cap (VC_ERR_CAPACITY as above);
 *   keep_edge   for each old edge, 1 if it survived as an edge (its measure is
 *               kept), else 0.
 * Validity flags as in vc_voro_extract_inside (only used with only_inside). */
int vc_voro_merge_close(int32_t nv, int32_t ne, int32_t nf, const float* vts,
                        const int32_t* edges, const int64_t* face_off,
                        const int32_t* face_vts, const int64_t* ve_off,
                        const int32_t* ve, const uint8_t* v_valid, float eps,
                        int only_inside, int32_t* old_to_new, int32_t* new_vts,
                        int32_t* n_new_vts, int64_t* new_face_off,
                        int32_t* new_face_vts, uint8_t* remove_face,
                        int32_t* new_edges, int64_t edge_cap,
                        int64_t* n_new_edges, uint8_t* keep_edge);

/* vc_voro_merge_close with the work on the device when no edge joins two
 * vertices within eps (no BFS absorbs anything: every eligible vertex is its
 * own cluster, faces keep their distinct new ids, the new edge set is sorted
 * and deduplicated on the device); otherwise the order-defined BFS runs on
 * the host (vc_voro_merge_close).  Same arguments and outputs; *on_device
 * (may be NULL) is 1 when the device path ran.  Waits. */
int vc_voro_merge_close_device(vc_ctx* ctx, int32_t nv, int32_t ne, int32_t nf,
                               const float* vts, const int32_t* edges,
                               const int64_t* face_off, const int32_t* face_vts,
                               const int64_t* ve_off, const int32_t* ve,
                               const uint8_t* v_valid, float eps, int only_inside,
                               int32_t* old_to_new, int32_t* new_vts,
                               int32_t* n_new_vts, int64_t* new_face_off,
                               int32_t* new_face_vts, uint8_t* remove_face,
                               int32_t* new_edges, int64_t edge_cap,
                               int64_t* n_new_edges, uint8_t* keep_edge,
                               int* on_device);

/* Radius of every vertex (host arrays, computed on the device, waits): faces
 * are visited in order and every vertex v of face f is given
 * |v - site[face_sites[2f]]| (float, ((d0^2 + d1^2) + d2^2), correctly
 * rounded sqrt), so the last face holding v decides; 0 for a vertex in no
 * face. */
int vc_voro_radii(vc_ctx* ctx, int32_t nv, int32_t nf, const float* vts,
                  const int64_t* face_off, const int32_t* face_vts,
                  const float* sites, int32_t nsites, const int32_t* face_sites,
                  float* radii);

/* Inside part (host): vts_idx = valid vertex ids ascending; tris = for every
 * face in order, the triangles of its zig-zag triangulation whose three
 * corners are valid (the face need not be), renumbered into vts_idx order;
 * edges = the sides of those triangles, makeEdge-normalised, deduplicated
 * and sorted ascending (the reference collects them in a std::set).
 * Capacities and VC_ERR_CAPACITY as in vc_voro_extract_inside. */
int vc_voro_part_inside(int32_t nv, int32_t nf, const int64_t* face_off,
                        const int32_t* face_vts, const uint8_t* v_valid,
                        int32_t* vts_idx, int64_t vts_cap, int64_t* n_vts,
                        int32_t* edges_out, int64_t edge_cap, int64_t* n_edges,
                        uint32_t* tris, int64_t tri_cap, int64_t* n_tris);

/* Discrete Voronoi diagram of the background voxels, restricted to the
 * object, from the feature transform (vc_edt must have run on vol): the
 * voxel-core construction that replaces TetGen in computeVD
 * (src/highlevelalgo.cpp:487-529, src/voroinfo.cpp:113-325), whose output
 * feeds the steps above unchanged.  Every pair of 6-adjacent inside voxels
 * p, q = p + e_a whose nearest background voxels (feature transform, lowest
 * linear index among the nearest) differ contributes the voxel face between
 * them, a quad on the bisector region of the two sites.
 *   faces       4 int32 vertex ids per face; faces ordered by p (linear
 *               index), then axis x, y, z; corners (p_b + db, p_c + dc) of
 *               the face at p_a + 1/2, b = a+1, c = a+2 (mod 3), in the order
 *               (0,0), (1,0), (1,1), (0,1): counter-clockwise seen from +a;
 *   face_sites  2 int32 site ids per face: p's site, then q's;
 *   vts         3 floats per vertex: the face corners used, numbered in
 *               corner-lattice order ((i, j, k) at (i-1/2, j-1/2, k-1/2),
 *               i slowest), in voxel coordinates;
 *   sites       3 floats per site: the background voxels that are a face
 *               site, numbered in linear order (voxel centres).
 * Lambda of a face is |s1 - s2| (vc_voro_measures).  Host outputs; with any
 * of vts / faces / face_sites / sites NULL only the counts are written (size
 * query); VC_ERR_CAPACITY with the counts written when a capacity is short;
 * VC_ERR_STATE before vc_edt; VC_ERR_INVALID when 3 * voxels reaches 2^32 or
 * the corner lattice 2^31 points.  Runs on the device; waits. */
int vc_voro_from_ft(vc_ctx* ctx, vc_vol* vol, float* vts, int64_t vts_cap,
                    int64_t* n_vts, int32_t* faces, int32_t* face_sites,
                    int64_t face_cap, int64_t* n_faces, float* sites,
                    int64_t site_cap, int64_t* n_sites);

#ifdef __cplusplus
}
#endif
#endif /* VC_VORO_H */
