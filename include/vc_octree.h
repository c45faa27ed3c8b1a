/*
 * vc_octree.h -- sparse octree volumes (.sof / .sog) as input of the path
 * (libvoxelcore.so).
 *
 * Reference entry points replaced (reference file:line):
 *   vc_octree_parse          OctreeVolume::read_sof_file / read_sog_file
 *                            (src/octree.cpp:259-497)
 *   vc_volume_from_octree    OctreeVolume::getDataAt for every voxel
 *                            (src/octree.cpp:200-255, get_node_values
 *                            :676-774) + SpaceConverter::get_occupancy_at_vox
 *                            (include/spaceinfo.h:122-135)
 *   vc_octree_leaf_centers   Surfacer::extractBoundaryVts, octree branch
 *                            (src/surfacing.cpp:286-314, :666-686)
 *
 * Stream: int32 resolution R (.sof: first 4 bytes; .sog: after a 128-byte
 * header "SOG.Format 1.0\0" + lower-left corner (3 float32) + bounding-box
 * edge (float32)), then the nodes in pre-order: type byte (0 internal,
 * 1 empty, 2 leaf); empty: one sign byte (1 = outside); leaf: one byte of the
 * 8 corner signs (.sog: followed by 3 float32, skipped); internal: its 8
 * children, (i, j, k) with k fastest.  R must be a power of two; the volume
 * has (R+1)^3 voxels.
 */
#ifndef VC_OCTREE_H
#define VC_OCTREE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vc_ctx vc_ctx;
typedef struct vc_vol vc_vol;
typedef struct vc_octree vc_octree;

#define VC_OCTREE_SOF 0
#define VC_OCTREE_SOG 1

typedef struct vc_octree_info {
    int32_t resolution;          /* R                                      */
    int64_t n_internal, n_empty, n_leaf;
    float lower_left[3];         /* .sog only (0 for .sof)                 */
    float bbox_len;              /* .sog only (0 for .sof)                 */
} vc_octree_info;

/* parse a whole file held in memory; VC_ERR_INVALID on a malformed stream */
int vc_octree_parse(const uint8_t* bytes, size_t n, int format,
                    vc_octree** out, vc_octree_info* info);
int vc_octree_free(vc_octree* t);
/* OctreeVolume::getDataAt on the host: +1 inside, -1 outside, for
 * 0 <= x, y, z <= R (a coordinate equal to R reads the neighbouring corner
 * of the cell below it) */
int vc_octree_value(const vc_octree* t, int x, int y, int z, double* value);
/* dense (R+1)^3 occupancy built on the device (one thread per voxel walks
 * the tree), as a volume of this context */
int vc_volume_from_octree(vc_ctx* ctx, const vc_octree* t, vc_vol** out);
/* centres (node corner + 0.5) of the leaves whose sign byte is neither 0 nor
 * 0xFF, in pre-order; VC_ERR_CAPACITY with *n_out set when cap is short */
int vc_octree_leaf_centers(const vc_octree* t, float* xyz, int64_t cap,
                           int64_t* n_out);

#ifdef __cplusplus
}
#endif
#endif /* VC_OCTREE_H */
