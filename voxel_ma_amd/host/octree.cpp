// octree.cpp -- OctreeVolume over include/vc_octree.h (see octree.h).
#include "octree.h"

#include <fstream>
#include <iterator>
#include <stdexcept>

#include "../../include/vc_api.h"
#include "../../include/vc_octree.h"
#include "vc_device.h"

OctreeVolume::OctreeVolume(const char* file)
{
    const std::string f(file ? file : "");
    const bool sog = f.size() >= 4 && f.compare(f.size() - 4, 4, ".sog") == 0;
    if (!read(f, sog))
        throw std::runtime_error("cannot read octree file " + f);
}

OctreeVolume::~OctreeVolume() { vc_octree_free(m_tree); }

bool OctreeVolume::read_sof_file(const std::string& file) { return read(file, false); }
bool OctreeVolume::read_sog_file(const std::string& file) { return read(file, true); }

bool OctreeVolume::read(const std::string& file, bool sog)
{
    std::ifstream in(file, std::ios::binary);
    if (!in)
        return false;
    const std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(in)),
                                     std::istreambuf_iterator<char>());
    vc_octree* t = nullptr;
    vc_octree_info info{};
    if (vc_octree_parse(bytes.data(), bytes.size(),
                        sog ? VC_OCTREE_SOG : VC_OCTREE_SOF, &t, &info) != VC_OK)
        return false;
    vc_octree_free(m_tree);
    m_tree = t;
    m_max_res = info.resolution;
    m_sizes = ivec3(m_max_res + 1, m_max_res + 1, m_max_res + 1);
    m_lower_left = point(info.lower_left[0], info.lower_left[1], info.lower_left[2]);
    m_bbox_len = info.bbox_len;
    if (sog)
    {
        // voxel -> model: scale by bbox edge / R, then translate to the lower
        // left corner (src/octree.cpp:461-475)
        const double sc = (double)m_bbox_len / (double)m_max_res;
        m_vox_to_mesh = xform::trans(m_lower_left) * xform::scale(sc);
        const point neg(-m_lower_left[0], -m_lower_left[1], -m_lower_left[2]);
        m_mesh_to_vox = xform::scale(1.0 / sc) * xform::trans(neg);
    }
    else
    {
        m_vox_to_mesh = xform::identity();
        m_mesh_to_vox = xform::identity();
    }
    return true;
}

double OctreeVolume::getDataAt(int x, int y, int z) const
{
    if (!m_tree)
        throw std::runtime_error("OctreeVolume: no octree loaded");
    double v = 1.0;
    if (vc_octree_value(m_tree, x, y, z, &v) != VC_OK)
        return 1.0;
    return v;
}

std::vector<uint8_t> OctreeVolume::occupancy() const
{
    if (!m_tree)
        throw std::runtime_error("OctreeVolume: no octree loaded");
    vc_ctx* ctx = vcdev::context();
    vc_vol* v = nullptr;
    vcdev::check(vc_volume_from_octree(ctx, m_tree, &v), "vc_volume_from_octree");
    const size_t n = (size_t)(m_max_res + 1);
    std::vector<uint8_t> occ(n * n * n);
    const int rc = vc_get(ctx, v, VC_BUF_OCC, occ.data(), (int64_t)occ.size());
    vc_volume_release(ctx, v);
    vcdev::check(rc, "vc_get");
    return occ;
}

void OctreeVolume::leafCenters(std::vector<point>& out) const
{
    out.clear();
    if (!m_tree)
        return;
    int64_t n = 0;
    vc_octree_leaf_centers(m_tree, nullptr, 0, &n);
    std::vector<float> xyz(3 * (size_t)n + 3);
    vcdev::check(vc_octree_leaf_centers(m_tree, xyz.data(), n, &n),
                 "vc_octree_leaf_centers");
    for (int64_t i = 0; i < n; ++i)
        out.push_back(point(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
}
