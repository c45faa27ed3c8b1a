// exporters.h -- the reference's writers (include/exporters.h:22-102,
// namespace voxelvoro) over include/vc_export.h: PLY, .r radii, QMAT .ma,
// boundary mesh / points as OFF and TetGen .node, byte for byte as the
// reference writes them.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "Volume3DScalar.h"
#include "commondefs.h"
#include "voroinfo.h"

namespace voxelvoro
{

enum class ExportErrCode
{
    FAILURE,
    SUCCESS,
    INPUT_NOT_OPEN,
    OUTPUT_NOT_OPEN,
    INPUT_NOT_SUPPORTED,
    OUTPUT_NOT_SUPPORTED
};

// boundary faces (or only vertices) of the volume, in model space, to
// <base>.off; optionally the voxel-space vertices to <base>.node
// (src/exporters.cpp:94-220)
ExportErrCode writeVolumeAsBoundaryMesh(const std::shared_ptr<Volume3DScalar>& vol,
                                        const char* outfile_base, bool need_euler,
                                        bool write_node = false);
ExportErrCode writeVolumeAsBoundaryPts(const std::shared_ptr<Volume3DScalar>& vol,
                                       const char* outfile_base,
                                       bool write_node = false);
ExportErrCode writeToTetnodes(const std::vector<point>& vts, const char* outfile);

// src/exporters.cpp:244-289
ExportErrCode writeToDotma(const std::string& filename, const std::vector<point>& vts,
                           const std::vector<ivec2> all_edges,
                           const std::vector<uTriFace>& tri_faces,
                           const std::vector<float>& radii);
// src/exporters.cpp:290-339 (getPartInside + untransformed radii)
ExportErrCode writeInsideVoroToDotMA(const VoroInfo& voro, const char* filename);

// src/exporters.cpp:341-476: the measures and sites are accepted as the
// reference accepts them, and, as there, not written (its PLY writer
// declares only x y z, vertex1 vertex2 and vertex_indices)
ExportErrCode writeToPLY(const char* filename, const std::vector<point>& vts,
                         const std::vector<ivec2>& edges,
                         const std::vector<uTriFace>& tris,
                         const std::vector<float>& vts_msure,
                         const std::vector<float>& edges_msure,
                         const std::vector<float>& faces_msure,
                         bool write_sites = false,
                         const std::vector<point>* sites = nullptr,
                         const std::vector<ivec2>* face_sites_ids = nullptr);

// src/exporters.cpp:478-552: valid vertices (or the given ids) and their
// radii, both through mat
ExportErrCode writeRadiiToFile(const VoroInfo& voro, const char* r_file,
                               const xform& mat,
                               const std::vector<int>* ids = nullptr);

} // namespace voxelvoro
