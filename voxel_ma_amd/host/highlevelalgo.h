// highlevelalgo.h -- the orchestration entry points of the thinning path
// (reference include/highlevelalgo.h:28-71, namespace voxelvoro).
#pragma once

#include <memory>
#include <vector>

#include "Volume3DScalar.h"
#include "ccthin.h"
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
} // namespace voxelvoro
