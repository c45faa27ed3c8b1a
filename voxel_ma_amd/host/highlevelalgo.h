// highlevelalgo.h -- the orchestration entry points of the thinning path
// (reference include/highlevelalgo.h:28-71, namespace voxelvoro).
#pragma once

#include <memory>
#include <vector>

#include "Volume3DScalar.h"
#include "ccthin.h"
#include "exporters.h"
#include "voroinfo.h"
#include "voxelthin.h"

namespace voxelvoro
{
// number of closed pockets: collapse everything collapsible (all measures 0,
// thresholds 1) and count the face components left
// (src/highlevelalgo.cpp:993-1020)
int nClosedComponents(const cellcomplex& cc);

// longest side of the bounding box of the inside voxel centres, the voxel
// form of VoroInfo::getInsidePartSize (src/voroinfo.cpp:496)
float insidePartSize(const Volume3DScalar& vol);

// the measure-ordered pruning loop of exportInsideVoroMesh
// (src/highlevelalgo.cpp:825-836): thresh = float(insideSize * t) for every
// t in tt (ascending), one subfield-thinning run over all of them
std::vector<float> thresholdsFromTT(float inside_size, std::vector<double> tt);
void thinVolume(const std::shared_ptr<Volume3DScalar>& vol,
                const std::vector<double>& tt, VoxelThinning& thin,
                std::vector<float>* levels_out = nullptr);

// the Voronoi diagram of vol2ma (src/highlevelalgo.cpp:487-529): the
// discrete diagram of the volume's feature transform (vc_voro_from_ft) in
// place of TetGen's diagram of the boundary sites; radii computed as the
// reference's diagram loader does
void computeVD(const std::shared_ptr<Volume3DScalar>& vol, VoroInfo& voro);
// computeVD with the diagram copied to the host (VoroInfo::setGeometry), so
// the later steps run on host arrays (each through its own device call)
void computeVDHost(const std::shared_ptr<Volume3DScalar>& vol, VoroInfo& voro);

// src/highlevelalgo.cpp:530-583: tag the vertices, merge those closer than
// insideSize / max axis * 1e-6 (generateMeasure is a no-op on the
// unfinalized diagram there, so it is not repeated)
bool preprocessVoro(VoroInfo& voro, const std::shared_ptr<Volume3DScalar>& vol,
                    bool need_euler);

// src/highlevelalgo.cpp:658-961, .off and .ply outputs: full inside part
// (full_or_thin 1 or 2) and the FIFO-pruned complex for every t >= 0 of tt
// (full_or_thin 0 or 2), with .r radii and, with out_qmat, .ma files
ExportErrCode exportInsideVoroMesh(VoroInfo& voro,
                                   const std::shared_ptr<Volume3DScalar>& vol,
                                   const char* mesh_file, bool need_euler,
                                   bool collapse_degenerate_edges, bool inside_only,
                                   bool finite_only, int full_or_thin,
                                   bool write_attrib, bool out_qmat,
                                   const std::vector<float>& tt);
} // namespace voxelvoro
