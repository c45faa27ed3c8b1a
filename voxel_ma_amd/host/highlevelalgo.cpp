#include "highlevelalgo.h"

#include <algorithm>
#include <iostream>
#include <set>
#include <sstream>

#include "../../include/vc_export.h"
#include "../../include/vc_vdiag.h"
#include "../../include/vc_voro.h"
#include "geomalgo.h"
#include "vc_device.h"

namespace voxelvoro
{

int nClosedComponents(const cellcomplex& cc_in)
{
    cellcomplex cc = cc_in;
    cc.finalize();
    int64_t n = 0;
    const int rc = vc_cc_closed_components(vcdev::context(), cc.device().get(), &n);
    if (rc != VC_ERR_INVALID)
    {
        vcdev::check(rc, "vc_cc_closed_components");
        return (int)n;
    }
    // a complex the device collapse refuses (an edge with equal ends, a face
    // listing a side twice): the same steps through the host collapse
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

void computeVD(const std::shared_ptr<Volume3DScalar>& vol, VoroInfo& voro)
{
    vc_ctx* ctx = vcdev::context();
    const std::vector<uint8_t> occ = vol->occupancy();
    vc_vol* v = nullptr;
    vcdev::check(vc_volume_upload(ctx, occ.data(), vol->getSizeX(),
                                  vol->getSizeY(), vol->getSizeZ(), &v),
                 "vc_volume_upload");
    struct Release
    {
        vc_ctx* c;
        vc_vol* v;
        ~Release() { vc_volume_release(c, v); }
    } release{ctx, v};
    vcdev::check(vc_edt(ctx, v, 0), "vc_edt");
    // the diagram stays on the device; VoroInfo copies parts of it to the
    // host only where a caller asks for them
    vc_vdiag* d = nullptr;
    vcdev::check(vc_vdiag_from_ft(ctx, v, &d), "vc_vdiag_from_ft");
    voro.setDevice(std::shared_ptr<vc_vdiag>(d, [ctx](vc_vdiag* p) { vc_vdiag_release(ctx, p); }));
    vc_vdiag_info in;
    vcdev::check(vc_vdiag_info_get(d, &in), "vc_vdiag_info_get");
    std::cout << "discrete voronoi diagram: " << in.nv << " vts, " << in.nf
              << " faces, " << in.ns << " sites" << std::endl;
    voro.computeRadii();
}

void computeVDHost(const std::shared_ptr<Volume3DScalar>& vol, VoroInfo& voro)
{
    vc_ctx* ctx = vcdev::context();
    const std::vector<uint8_t> occ = vol->occupancy();
    vc_vol* v = nullptr;
    vcdev::check(vc_volume_upload(ctx, occ.data(), vol->getSizeX(),
                                  vol->getSizeY(), vol->getSizeZ(), &v),
                 "vc_volume_upload");
    struct Release
    {
        vc_ctx* c;
        vc_vol* v;
        ~Release() { vc_volume_release(c, v); }
    } release{ctx, v};
    vcdev::check(vc_edt(ctx, v, 0), "vc_edt");
    int64_t nv = 0, nf = 0, ns = 0;
    vcdev::check(vc_voro_from_ft(ctx, v, nullptr, 0, &nv, nullptr, nullptr, 0,
                                 &nf, nullptr, 0, &ns),
                 "vc_voro_from_ft");
    std::vector<float> vts(3 * (size_t)nv + 3), sites(3 * (size_t)ns + 3);
    std::vector<int32_t> faces(4 * (size_t)nf + 4), fs(2 * (size_t)nf + 2);
    vcdev::check(vc_voro_from_ft(ctx, v, vts.data(), nv, &nv, faces.data(),
                                 fs.data(), nf, &nf, sites.data(), ns, &ns),
                 "vc_voro_from_ft");
    std::vector<point> pv((size_t)nv), ps((size_t)ns);
    for (int64_t i = 0; i < nv; ++i)
        pv[i] = point(vts[3 * i], vts[3 * i + 1], vts[3 * i + 2]);
    for (int64_t i = 0; i < ns; ++i)
        ps[i] = point(sites[3 * i], sites[3 * i + 1], sites[3 * i + 2]);
    std::vector<std::vector<int>> pf((size_t)nf);
    std::vector<ivec2> pfs((size_t)nf);
    for (int64_t f = 0; f < nf; ++f)
    {
        pf[f].assign(faces.begin() + 4 * f, faces.begin() + 4 * f + 4);
        pfs[f] = ivec2(fs[2 * f], fs[2 * f + 1]);
    }
    std::cout << "discrete voronoi diagram: " << nv << " vts, " << nf
              << " faces, " << ns << " sites" << std::endl;
    voro.setGeometry(pv, pf, pfs, ps);
    voro.computeRadii();
}

bool preprocessVoro(VoroInfo& voro, const std::shared_ptr<Volume3DScalar>& vol,
                    bool need_euler)
{
    if (!voro.onlyHasInside() && !voro.tagVtsUsingUniformVol(vol))
    {
        std::cout << "Error: couldn't tag voro vts." << std::endl;
        return false;
    }
    eulerchar ec;
    if (need_euler)
    {
        voro.geom().eulerChar(ec);
        ec.logToConsole("pruned voro diagram");
    }
    const int max_len =
        std::max(std::max(vol->getSizeX(), vol->getSizeY()), vol->getSizeZ());
    const float eps = voro.getInsidePartSize() * (1.0 / max_len) * 0.000001;
    std::cout << "before merging voro size: " << voro.numVts() << "/"
              << voro.numEdges() << "/" << voro.numFaces()
              << std::endl;
    voro.mergeCloseVts(eps, true);
    std::cout << "after merging voro size: " << voro.numVts() << "/"
              << voro.numEdges() << "/" << voro.numFaces()
              << std::endl;
    if (need_euler)
    {
        voro.geom().eulerChar(ec);
        ec.logToConsole((std::string("after merging vts within ") +
                         std::to_string(eps))
                            .c_str());
    }
    return true;
}

ExportErrCode exportInsideVoroMesh(VoroInfo& voro,
                                   const std::shared_ptr<Volume3DScalar>& vol,
                                   const char* mesh_file_c, bool need_euler,
                                   bool /*collapse_degenerate_edges*/,
                                   bool inside_only, bool finite_only,
                                   int full_or_thin, bool write_attrib,
                                   bool out_qmat, const std::vector<float>& tt)
{
    const std::string mesh_file(mesh_file_c);
    const xform& mat = vol->getVoxToModelMat();
    if (mesh_file.find(".off") != std::string::npos)
    {
        // triangulated faces (valid, finite or all) over the model-space
        // vertices, unused vertices dropped in order
        std::vector<int> fids;
        if (inside_only)
            voro.getValidFaces(fids);
        else
            for (int f = 0; f < (int)voro.geom().numFaces(); ++f)
                if (!finite_only || voro.isFaceFinite(f))
                    fids.push_back(f);
        std::vector<uTriFace> tris;
        std::vector<int> poly;
        for (int f : fids)
        {
            voro.geom().getFaceVRep(f, poly);
            util::simpleTriangulate(poly, tris);
        }
        std::vector<int> remap(voro.geom().numVts(), -1);
        std::vector<point> vts;
        for (auto& t : tris)
            for (int k = 0; k < 3; ++k)
                if (remap[t[k]] < 0)
                    remap[t[k]] = 0;
        for (size_t i = 0; i < remap.size(); ++i)
            if (remap[i] == 0)
            {
                remap[i] = (int)vts.size();
                vts.push_back(mat * voro.geom().getVert(i));
            }
        std::vector<uint32_t> flat;
        for (auto& t : tris)
            for (int k = 0; k < 3; ++k)
                flat.push_back((uint32_t)remap[t[k]]);
        std::vector<float> fv;
        for (auto& p : vts)
            fv.insert(fv.end(), p.begin(), p.end());
        return vc_write_off(mesh_file_c, fv.data(), (int64_t)vts.size(),
                            flat.data(), (int64_t)tris.size(), nullptr) == VC_OK
                   ? ExportErrCode::SUCCESS
                   : ExportErrCode::FAILURE;
    }
    if (mesh_file.find(".ply") == std::string::npos)
        return ExportErrCode::OUTPUT_NOT_SUPPORTED;

    const std::string base = mesh_file.substr(0, mesh_file.rfind('.'));
    std::vector<point> vts;
    std::vector<ivec2> edges;
    std::vector<uTriFace> tris;
    std::vector<int> from_fi;
    std::vector<float> vm, em, fm;
    voro.extractInsideWithMeasure(MeasureForMA::LAMBDA, vts, edges, tris,
                                  from_fi, vm, em, fm);
    std::cout << "inside part extracted. # vts/edges/faces: "
              << voro.numVts() << "/" << voro.numEdges() << "/"
              << voro.numFaces() << std::endl;
    if (vts.empty())
    {
        std::cout << "IVD: 0 vts; IO skipped." << std::endl;
        return ExportErrCode::SUCCESS;
    }
    std::vector<point> vts_trans = vts;
    for (auto& p : vts_trans)
        p = mat * p;
    if (full_or_thin == 1 || full_or_thin == 2)
    {
        if (write_attrib)
        {
            std::vector<ivec2> site_ids;
            for (int f : from_fi)
                site_ids.push_back(voro.getSitesOfFace(f));
            writeToPLY(mesh_file_c, vts_trans, edges, tris, vm, em, fm, true,
                       &voro.getSitesPosition(), &site_ids);
        }
        else
            writeToPLY(mesh_file_c, vts_trans, edges, tris, vm, em, fm);
        writeRadiiToFile(voro, (base + ".r").c_str(), mat);
    }
    if (full_or_thin == 1)
        return ExportErrCode::SUCCESS;

    CellComplexThinning ccthin;
    cellcomplex cc(vts_trans, edges, tris, false);
    // the extraction on the device already finalized this complex there
    if (auto dev = voro.takeInsideComplex())
        cc.adoptFinalized(std::move(dev));
    else
        cc.finalize();
    ccthin.setup(&cc);
    ccthin.assignElementValues(vm, em, fm);
    ccthin.preprocess();
    const point orig = mat * point(0, 0, 0);
    for (double t : tt)
    {
        if (t < 0.0)
            continue;
        const float thresh = voro.getInsidePartSize() * t;
        ccthin.prune(thresh, thresh, false);
        ccthin.remainingCC(vts_trans, edges, tris, nullptr);
        if (need_euler)
        {
            eulerchar ec;
            util::computeEulerChar(vts_trans, edges, tris, ec);
            ec.logToConsole("voro diagram after thinning");
        }
        if (vts_trans.empty())
        {
            std::cout << "Thinned IVD: 0 vts remaining; IO skipped." << std::endl;
            return ExportErrCode::SUCCESS;
        }
        std::ostringstream ss;
        ss << t;
        const std::string end_str = std::string("_thinned") + ss.str();
        std::string thin_file = mesh_file;
        thin_file.insert(thin_file.find(".ply"), end_str);
        if (write_attrib)
        {
            std::vector<int> rf;
            ccthin.getRemainedFaces(rf);
            std::vector<ivec2> site_ids;
            std::vector<float> fm_thin;
            for (int ti : rf)
            {
                site_ids.push_back(voro.getSitesOfFace(from_fi[ti]));
                fm_thin.push_back(fm[ti]);
            }
            writeToPLY(thin_file.c_str(), vts_trans, edges, tris, {}, {},
                       fm_thin, true, &voro.getSitesPosition(), &site_ids);
        }
        else
            writeToPLY(thin_file.c_str(), vts_trans, edges, tris, {}, {}, {},
                       false);
        std::vector<int> remain;
        ccthin.getRemainedVts(remain);
        writeRadiiToFile(voro, (base + end_str + ".r").c_str(), mat, &remain);
        if (out_qmat)
        {
            // radii of the remaining ids as the reference takes them (the
            // diagram's per-vertex radii), through the transform
            const auto& R = voro.getRadii();
            std::vector<float> radii;
            for (int i : remain)
            {
                const point q = mat * point(i < (int)R.size() ? R[i] : 0.0f, 0, 0);
                const float d0 = q[0] - orig[0], d1 = q[1] - orig[1],
                            d2 = q[2] - orig[2];
                radii.push_back(std::sqrt(d0 * d0 + d1 * d1 + d2 * d2));
            }
            // all triangle sides, deduplicated, ascending (the reference
            // appends an unordered_set, whose order is implementation-defined)
            // (as packed (min, max) keys, sorted and made unique)
            std::vector<uint64_t> sides;
            sides.reserve(3 * tris.size());
            for (const auto& f : tris)
                for (int k = 0; k < 3; ++k)
                {
                    const uint32_t a = f[k], b = f[(k + 1) % 3];
                    sides.push_back((uint64_t)std::min(a, b) << 32 | std::max(a, b));
                }
            std::sort(sides.begin(), sides.end());
            sides.erase(std::unique(sides.begin(), sides.end()), sides.end());
            std::vector<ivec2> all_edges = edges;
            all_edges.reserve(edges.size() + sides.size());
            for (const uint64_t k : sides)
                all_edges.push_back(ivec2((int)(k >> 32), (int)(k & 0xFFFFFFFFu)));
            writeToDotma(thin_file.substr(0, thin_file.find_last_of('.')) + ".ma",
                         vts_trans, all_edges, tris, radii);
        }
    }
    return ExportErrCode::SUCCESS;
}

} // namespace voxelvoro
