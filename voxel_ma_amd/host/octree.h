// octree.h -- OctreeVolume with the reference's names (include/octree.h,
// src/octree.cpp:111-497): a .sof / .sog sparse octree as a volume.
// getDataAt follows OctreeVolume::getDataAt; occupancy() builds the dense
// (R+1)^3 grid on the device (include/vc_octree.h).
#pragma once

#include <string>
#include <vector>

#include "Volume3DScalar.h"
#include "commondefs.h"

struct vc_octree;

class OctreeVolume : public Volume3DScalar
{
public:
    OctreeVolume() = default;
    // reads a .sof or .sog file (by extension); throws std::runtime_error
    explicit OctreeVolume(const char* file);
    ~OctreeVolume() override;
    OctreeVolume(const OctreeVolume&) = delete;
    OctreeVolume& operator=(const OctreeVolume&) = delete;

    bool read_sof_file(const std::string& file);
    bool read_sog_file(const std::string& file);
    // +1 inside, -1 outside (src/octree.cpp:200-255); coordinates outside
    // [0, R] read no node and give +1, as the reference's lookup does
    double getDataAt(int x, int y, int z) const override;
    std::vector<uint8_t> occupancy() const override;
    int maxRes() const { return m_max_res; }
    // centres of the leaves with mixed signs, in tree order
    // (Surfacer::OctreeLeafCenterCollector, src/surfacing.cpp:666-686)
    void leafCenters(std::vector<point>& out) const;
    // .sog only: lower-left corner and bounding-box edge in model space
    const point& lowerLeft() const { return m_lower_left; }
    float bboxLen() const { return m_bbox_len; }

private:
    bool read(const std::string& file, bool sog);
    vc_octree* m_tree = nullptr;
    int m_max_res = 0;
    point m_lower_left;
    float m_bbox_len = 0.0f;
};
