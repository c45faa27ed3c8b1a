// voxelthin.h -- VoxelThinning: the voxel counterpart of
// CellComplexThinning (same setup / assign / prune / getRemained* shape,
// include/ccthin.h:57-113) over the GPU path: exact squared EDT + feature
// transform, per-voxel lambda measure, measure-ordered subfield thinning.
#pragma once

#include <memory>
#include <vector>

#include "Volume3DScalar.h"
#include "commondefs.h"
#include "../../include/vc_api.h"

class VoxelThinning
{
public:
    VoxelThinning() = default;
    ~VoxelThinning() { reset(); }
    VoxelThinning(const VoxelThinning&) = delete;
    VoxelThinning& operator=(const VoxelThinning&) = delete;

    void setup(const std::shared_ptr<Volume3DScalar>& vol); // upload to HBM
    void reset();
    // exact measure: EDT/FT, then lambda^2 per voxel
    void computeMeasure();
    // caller-supplied per-voxel measure (C order [x][y][z]), e.g. a surrogate
    void assignMeasure(const std::vector<float>& measure);
    // levels ascending; see vc_thin_levels
    void prune(const std::vector<float>& levels);
    void prune(float level) { prune(std::vector<float>{level}); }
    // skeleton voxels remaining after level k
    void getRemainedVoxels(int k, std::vector<ivec3>& vox) const;
    const std::vector<uint8_t>& labels() const { return m_labels; }
    std::vector<int32_t> squaredDistances() const;
    std::vector<uint32_t> lambda2() const;
    const vc_thin_stats& stats() const { return m_stats; }
    int sizeX() const { return m_n[0]; }
    int sizeY() const { return m_n[1]; }
    int sizeZ() const { return m_n[2]; }

private:
    vc_vol* m_vol = nullptr;
    int m_n[3] = {0, 0, 0};
    bool m_has_measure = false;
    std::vector<uint8_t> m_labels;
    vc_thin_stats m_stats{};
};
