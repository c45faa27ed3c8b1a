#include "highlevelalgo.h"

#include <algorithm>

namespace voxelvoro
{

int nClosedComponents(const cellcomplex& cc_in)
{
    cellcomplex cc = cc_in;
    cc.finalize();
    CellComplexThinning thin;
    thin.setup(&cc);
    thin.assignElementValues(std::vector<float>(cc.numVts(), 0.0f),
                             std::vector<float>(cc.numEdges(), 0.0f),
                             std::vector<float>(cc.numFaces(), 0.0f));
    thin.preprocess();
    thin.prune(1.0f, 1.0f, false);
    cellcomplex remain = thin.remainingCC();
    if (remain.numFaces() == 0)
        return 0;
    return remain.compute_conn_cmpnts();
}

float insidePartSize(const Volume3DScalar& vol)
{
    int lo[3] = {vol.getSizeX(), vol.getSizeY(), vol.getSizeZ()};
    int hi[3] = {-1, -1, -1};
    const std::vector<uint8_t> occ = vol.occupancy();
    size_t i = 0;
    for (int x = 0; x < vol.getSizeX(); ++x)
        for (int y = 0; y < vol.getSizeY(); ++y)
            for (int z = 0; z < vol.getSizeZ(); ++z, ++i)
                if (occ[i])
                {
                    const int p[3] = {x, y, z};
                    for (int a = 0; a < 3; ++a)
                    {
                        lo[a] = std::min(lo[a], p[a]);
                        hi[a] = std::max(hi[a], p[a]);
                    }
                }
    int size = 0;
    for (int a = 0; a < 3; ++a)
        if (hi[a] >= lo[a])
            size = std::max(size, hi[a] - lo[a]);
    return (float)size;
}

std::vector<float> thresholdsFromTT(float inside_size, std::vector<double> tt)
{
    std::sort(tt.begin(), tt.end());
    std::vector<float> lv;
    for (double t : tt)
        if (t >= 0.0)
            lv.push_back((float)(inside_size * t));
    return lv;
}

void thinVolume(const std::shared_ptr<Volume3DScalar>& vol,
                const std::vector<double>& tt, VoxelThinning& thin,
                std::vector<float>* levels_out)
{
    const std::vector<float> lv = thresholdsFromTT(insidePartSize(*vol), tt);
    thin.setup(vol);
    thin.computeMeasure();
    thin.prune(lv);
    if (levels_out)
        *levels_out = lv;
}

} // namespace voxelvoro
