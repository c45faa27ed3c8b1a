/*
 * vc_vdiag.h -- the discrete Voronoi diagram of a volume kept resident on the
 * device through the voxel-core steps of vol2ma (libvoxelcore.so): built
 * from the feature transform, given radii, tagged, merged, and its inside
 * part extracted with measures into a resident complex for the collapse, so
 * no step copies the diagram between host and device.  VoroInfo holds one
 * after computeVD and copies parts of it to the host only when a caller asks
 * for them.
 *
 * The diagram is a finalized 2-complex (the cellcomplex numbering of
 * include/vc_cc.h: edges are the face sides in face order) with the two
 * sites of every face, the site positions, a validity flag per vertex
 * (cleared until tagged) and, once computed, a radius per vertex.
 *
 * Reference entry points replaced (reference file:line):
 *   vc_vdiag_from_ft      computeVD's diagram (src/highlevelalgo.cpp:487-529)
 *                         as VoroInfo geometry, the discrete diagram of
 *                         vc_voro_from_ft (include/vc_voro.h)
 *   vc_vdiag_radii        the radius step of VoroInfo::computeInfoRelatedtoSites
 *                         (src/voroinfo.cpp:290-325), as vc_voro_radii
 *   vc_vdiag_tag          VoroInfo::tagVtsUsingUniformVol (src/voroinfo.cpp:
 *                         447-494) and the valid-vertex box behind
 *                         getInsidePartSize, as vc_voro_tag_points
 *   vc_vdiag_merge_close  VoroInfo::mergeCloseVts (src/voroinfo.cpp:1064-1384)
 *                         when no edge joins two vertices within eps, as
 *                         vc_voro_merge_close_device
 *   vc_vdiag_extract_inside  VoroInfo::extractInsideWithMeasure with LAMBDA
 *                         (src/voroinfo.cpp:624-729, :1432-1574), as
 *                         vc_voro_measures + vc_voro_extract_inside_device,
 *                         and the cellcomplex exportInsideVoroMesh builds from
 *                         its output (src/highlevelalgo.cpp:819-826)
 *   vc_vdiag_get          the diagram's arrays on the host
 */
#ifndef VC_VDIAG_H
#define VC_VDIAG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vc_ctx vc_ctx;
typedef struct vc_vol vc_vol;
typedef struct vc_cc vc_cc;
typedef struct vc_vdiag vc_vdiag;

typedef struct vc_vdiag_info {
    int64_t nv, ne, nf, ns; /* vertices, edges, faces, sites */
    int64_t n_sides;        /* total face sides (face_vts entries) */
    int32_t tagged, radii;  /* 1 once vc_vdiag_tag / vc_vdiag_radii ran */
    int64_t n_valid;        /* vertices flagged valid */
    float bb_min[3], bb_max[3]; /* box of the valid vertices at tagging
                                   (voxel coordinates); zeros when none */
} vc_vdiag_info;

/* The diagram of vc_voro_from_ft (vc_edt must have run on vol), finalized;
 * no vertex valid yet.  Waits. */
int vc_vdiag_from_ft(vc_ctx* ctx, vc_vol* vol, vc_vdiag** out);
int vc_vdiag_release(vc_ctx* ctx, vc_vdiag* d);
int vc_vdiag_info_get(const vc_vdiag* d, vc_vdiag_info* info);
/* radius of every vertex as vc_voro_radii (the last face holding it
 * decides); kept with the diagram and carried through merges.  Waits. */
int vc_vdiag_radii(vc_ctx* ctx, vc_vdiag* d);
/* valid[v] = vertex v rounds to an inside voxel of vol (vc_voro_tag_points);
 * also the box of the valid vertices.  Waits. */
int vc_vdiag_tag(vc_ctx* ctx, vc_vdiag* d, vc_vol* vol);
/* mergeCloseVts(eps, only_inside) on the device when no edge joins two
 * vertices within eps (*on_device = 1): the diagram becomes the merged one
 * (seeds in order with their validity and radius, kept faces with their
 * sites, the sorted new edges, finalized again).  Otherwise the diagram is
 * unchanged and *on_device = 0: the order-defined merge runs on the host
 * (vc_voro_merge_close on vc_vdiag_get's arrays).  Waits. */
int vc_vdiag_merge_close(vc_ctx* ctx, vc_vdiag* d, float eps, int only_inside,
                         int* on_device);
/* Host copies; any output may be NULL: vts (3 nv floats), edges (2 ne),
 * face_off (nf + 1), face_vts (n_sides), face_sites (2 nf), sites (3 ns),
 * radii (nv; VC_ERR_STATE before vc_vdiag_radii), valid (nv bytes).
 * Waits. */
int vc_vdiag_get(vc_ctx* ctx, vc_vdiag* d, float* vts, int32_t* edges,
                 int64_t* face_off, int32_t* face_vts, int32_t* face_sites,
                 float* sites, float* radii, uint8_t* valid);
/* extractInsideWithMeasure(LAMBDA) with vertex validity valid[v] (or every
 * vertex with all_valid): the outputs of vc_voro_extract_inside_device plus
 * vm[i] = lambda of vertex vts_idx[i] (vc_voro_measures over the valid
 * faces).  Capacities: vts_cap >= nv, tri_cap >= n_sides - 2 nf and edge_cap
 * >= ne + 3 tri_cap always suffice; VC_ERR_CAPACITY with the counts written
 * when one is short.  With inside != NULL the complex of the output (vertices
 * 0..n_vts-1, the edges as given, the triangles as faces) is also created,
 * finalized, on the device (release it with vc_cc_release).  Waits. */
int vc_vdiag_extract_inside(vc_ctx* ctx, vc_vdiag* d, int all_valid,
                            int32_t* vts_idx, float* vm, int64_t vts_cap,
                            int64_t* n_vts, int32_t* edges, float* em,
                            int64_t edge_cap, int64_t* n_edges, int32_t* tris,
                            int32_t* from_fi, float* fm, int64_t tri_cap,
                            int64_t* n_tris, vc_cc** inside);

#ifdef __cplusplus
}
#endif
#endif /* VC_VDIAG_H */
