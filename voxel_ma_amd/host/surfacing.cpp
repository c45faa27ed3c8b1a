#include "surfacing.h"

#include <memory>

#include "octree.h"

#include "vc_device.h"

SurfacerErrCode Surfacer::run(const std::shared_ptr<Volume3DScalar>& vol,
                              std::vector<point>* vts, std::vector<ivec4>* quads,
                              eulerchar* euler) const
{
    if (!vol)
        return SurfacerErrCode::EMPTY_VOL;
    if (vol->getSizeX() <= 0 || vol->getSizeY() <= 0 || vol->getSizeZ() <= 0)
        return SurfacerErrCode::INVALID_VOL_TYPE;
    vc_ctx* ctx = vcdev::context();
    const std::vector<uint8_t> occ = vol->occupancy();
    vc_vol* v = nullptr;
    if (vc_volume_upload(ctx, occ.data(), vol->getSizeX(), vol->getSizeY(),
                         vol->getSizeZ(), &v) != VC_OK)
        return SurfacerErrCode::FAILURE;
    std::unique_ptr<vc_vol, void (*)(vc_vol*)> guard(
        v, [](vc_vol* p) { vc_volume_release(vcdev::context(), p); });
    int64_t ns = 0, nq = 0;
    vc_euler eu{};
    int rc = vc_boundary_extract(ctx, v, nullptr, 0, &ns, nullptr, 0, &nq, &eu);
    if (rc != VC_OK && rc != VC_ERR_CAPACITY)
        return SurfacerErrCode::FAILURE;
    std::vector<float> s(3 * (size_t)ns);
    std::vector<int32_t> q(4 * (size_t)nq);
    rc = vc_boundary_extract(ctx, v, s.data(), ns, &ns, q.data(), nq, &nq, &eu);
    if (rc != VC_OK)
        return SurfacerErrCode::FAILURE;
    if (vts)
    {
        vts->resize((size_t)ns);
        for (int64_t i = 0; i < ns; ++i)
            (*vts)[i] = point(s[3 * i], s[3 * i + 1], s[3 * i + 2]);
    }
    if (quads)
    {
        quads->resize((size_t)nq);
        for (int64_t i = 0; i < nq; ++i)
            (*quads)[i] = ivec4(q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]);
    }
    if (euler)
        *euler = eulerchar((int)eu.V, (int)eu.E, (int)eu.F, (int)eu.T,
                           (int)eu.C, (int)eu.euler);
    return SurfacerErrCode::SUCCESS;
}

SurfacerErrCode Surfacer::extractBoundaryVtsAndFaces(
    const std::shared_ptr<Volume3DScalar>& vol, std::vector<point>& vts,
    std::vector<ivec4>& faces, bool)
{
    return run(vol, &vts, &faces, nullptr);
}

SurfacerErrCode Surfacer::extractBoundaryVtsAndFaces(
    const std::shared_ptr<Volume3DScalar>& vol, std::vector<point>& vts,
    std::vector<uTriFace>& faces, eulerchar* euler)
{
    std::vector<ivec4> quads;
    const SurfacerErrCode rc = run(vol, &vts, &quads, euler);
    if (rc != SurfacerErrCode::SUCCESS)
        return rc;
    faces.clear();
    faces.reserve(2 * quads.size());
    for (const auto& f : quads)
    {
        faces.push_back(uTriFace(f[0], f[1], f[2]));
        faces.push_back(uTriFace(f[2], f[3], f[0]));
    }
    return SurfacerErrCode::SUCCESS;
}

SurfacerErrCode Surfacer::extractBoundaryVts(const std::shared_ptr<Volume3DScalar>& vol,
                                             std::vector<point>& vts)
{
    // octree input: the centres of the mixed-sign leaves, in tree order
    // (src/surfacing.cpp:286-314)
    if (auto oct = std::dynamic_pointer_cast<OctreeVolume>(vol))
    {
        oct->leafCenters(vts);
        return SurfacerErrCode::SUCCESS;
    }
    return run(vol, &vts, nullptr, nullptr);
}

int Surfacer::computeEulerChar(const std::vector<point>&, const std::vector<ivec4>&,
                               const std::shared_ptr<Volume3DScalar>& vol,
                               eulerchar& euler) const
{
    if (run(vol, nullptr, nullptr, &euler) != SurfacerErrCode::SUCCESS)
        return 0;
    return euler.euler;
}
