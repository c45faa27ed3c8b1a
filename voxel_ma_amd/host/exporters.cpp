// exporters.cpp -- voxelvoro writers over include/vc_export.h (see
// exporters.h).
#include "exporters.h"

#include <iostream>

#include "../../include/vc_api.h"
#include "../../include/vc_export.h"
#include "surfacing.h"

namespace voxelvoro
{
namespace
{

std::vector<float> flat(const std::vector<point>& p)
{
    std::vector<float> f(3 * p.size() + 3);
    for (size_t i = 0; i < p.size(); ++i)
        for (int a = 0; a < 3; ++a)
            f[3 * i + a] = p[i][a];
    return f;
}

template <class T, int N, class U>
std::vector<T> flat_idx(const std::vector<VcVec<N, U>>& e)
{
    std::vector<T> f(N * e.size() + N);
    for (size_t i = 0; i < e.size(); ++i)
        for (int a = 0; a < N; ++a)
            f[N * i + a] = (T)e[i][a];
    return f;
}

ExportErrCode status(int rc, const char* what, const char* file)
{
    if (rc == VC_OK)
        return ExportErrCode::SUCCESS;
    std::cout << "Error: " << what << " failed for " << (file ? file : "(null)")
              << " (" << rc << ")" << std::endl;
    return rc == VC_ERR_IO ? ExportErrCode::OUTPUT_NOT_OPEN : ExportErrCode::FAILURE;
}

} // namespace

ExportErrCode writeToTetnodes(const std::vector<point>& vts, const char* outfile)
{
    const auto v = flat(vts);
    return status(vc_write_tetnodes(outfile, v.data(), (int64_t)vts.size()),
                  "writeToTetnodes", outfile);
}

ExportErrCode writeVolumeAsBoundaryMesh(const std::shared_ptr<Volume3DScalar>& vol,
                                        const char* outfile_base, bool need_euler,
                                        bool write_node)
{
    Surfacer surf;
    std::vector<point> vts;
    std::vector<uTriFace> faces;
    eulerchar euler;
    const auto rc = surf.extractBoundaryVtsAndFaces(vol, vts, faces,
                                                    need_euler ? &euler : nullptr);
    if (rc == SurfacerErrCode::EMPTY_VOL)
        std::cout << "Error: empty volume was given!" << std::endl;
    if (rc != SurfacerErrCode::SUCCESS)
        return ExportErrCode::FAILURE;
    if (need_euler)
        euler.logToConsole("boundary voxel mesh");
    const std::string base(outfile_base);
    const auto v = flat(vts);
    const auto t = flat_idx<uint32_t>(faces);
    auto st = status(vc_write_off((base + ".off").c_str(), v.data(), (int64_t)vts.size(),
                                  t.data(), (int64_t)faces.size(),
                                  vol->getVoxToModelMat().data()),
                     "write .off", outfile_base);
    if (write_node)
    {
        const auto sn = writeToTetnodes(vts, (base + ".node").c_str());
        if (st == ExportErrCode::SUCCESS)
            st = sn;
    }
    return st;
}

ExportErrCode writeVolumeAsBoundaryPts(const std::shared_ptr<Volume3DScalar>& vol,
                                       const char* outfile_base, bool write_node)
{
    Surfacer surf;
    std::vector<point> vts;
    const auto rc = surf.extractBoundaryVts(vol, vts);
    if (rc == SurfacerErrCode::EMPTY_VOL)
        std::cout << "Error: given volume is empty!" << std::endl;
    if (rc != SurfacerErrCode::SUCCESS)
        return ExportErrCode::FAILURE;
    const std::string base(outfile_base);
    const auto v = flat(vts);
    const auto st = status(vc_write_off((base + ".off").c_str(), v.data(),
                                        (int64_t)vts.size(), nullptr, 0,
                                        vol->getVoxToModelMat().data()),
                           "write .off", outfile_base);
    if (write_node)
    {
        const auto sn = writeToTetnodes(vts, (base + ".node").c_str());
        if (sn != ExportErrCode::SUCCESS)
            return sn;
    }
    return st;
}

ExportErrCode writeToDotma(const std::string& filename, const std::vector<point>& vts,
                           const std::vector<ivec2> all_edges,
                           const std::vector<uTriFace>& tri_faces,
                           const std::vector<float>& radii)
{
    if (radii.size() < vts.size())
        return ExportErrCode::FAILURE;
    const auto v = flat(vts);
    const auto e = flat_idx<int32_t>(all_edges);
    const auto t = flat_idx<uint32_t>(tri_faces);
    return status(vc_write_dotma(filename.c_str(), v.data(), radii.data(),
                                 (int64_t)vts.size(), e.data(),
                                 (int64_t)all_edges.size(), t.data(),
                                 (int64_t)tri_faces.size()),
                  "writeToDotma", filename.c_str());
}

ExportErrCode writeInsideVoroToDotMA(const VoroInfo& voro, const char* filename)
{
    std::vector<int> vi;
    std::vector<ivec2> edges;
    std::vector<uTriFace> tris;
    voro.getPartInside(vi, edges, tris);
    const auto& r = voro.getRadii();
    std::vector<point> vts;
    std::vector<float> radii;
    for (int i : vi)
    {
        vts.push_back(voro.vertex(i));
        radii.push_back(i < (int)r.size() ? r[i] : 0.0f);
    }
    return writeToDotma(filename, vts, edges, tris, radii);
}

ExportErrCode writeToPLY(const char* filename, const std::vector<point>& vts,
                         const std::vector<ivec2>& edges,
                         const std::vector<uTriFace>& tris,
                         const std::vector<float>&, const std::vector<float>&,
                         const std::vector<float>&, bool, const std::vector<point>*,
                         const std::vector<ivec2>*)
{
    const auto v = flat(vts);
    const auto e = flat_idx<int32_t>(edges);
    std::vector<int32_t> t(3 * tris.size() + 3);
    for (size_t i = 0; i < tris.size(); ++i)
        for (int a = 0; a < 3; ++a)
            t[3 * i + a] = (int32_t)tris[i][a];
    std::cout << "writing to ply file... " << std::endl;
    std::cout << "# vts/edges/faces: " << vts.size() << "/" << edges.size() << "/"
              << tris.size() << std::endl;
    return status(vc_write_ply(filename, v.data(), (int64_t)vts.size(), e.data(),
                               (int64_t)edges.size(), t.data(), (int64_t)tris.size()),
                  "writeToPLY", filename);
}

ExportErrCode writeRadiiToFile(const VoroInfo& voro, const char* r_file,
                               const xform& mat, const std::vector<int>* ids)
{
    if (!voro.radiiValid())
    {
        std::cout << "Nothing written to file. Radii of the given voro is not "
                     "correctly set yet!"
                  << std::endl;
        return ExportErrCode::FAILURE;
    }
    std::vector<int> sel;
    if (ids)
        sel = *ids;
    else
        voro.getValidVts(sel);
    const auto& R = voro.getRadii();
    std::vector<float> v(3 * sel.size() + 3), r(sel.size() + 1);
    for (size_t k = 0; k < sel.size(); ++k)
    {
        const point& p = voro.vertex(sel[k]);
        for (int a = 0; a < 3; ++a)
            v[3 * k + a] = p[a];
        r[k] = R[sel[k]];
    }
    const auto st = status(vc_write_radii(r_file, v.data(), r.data(),
                                          (int64_t)sel.size(), mat.data()),
                           "writeRadiiToFile", r_file);
    if (st == ExportErrCode::SUCCESS)
        std::cout << "Done: " << sel.size() << " radii written to file " << r_file
                  << std::endl;
    return st;
}

} // namespace voxelvoro
