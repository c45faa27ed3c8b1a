// voroUtility.cpp -- command-line front end keeping the reference's mode and
// flag names (src/voroUtility.cpp:21-142) for the modes of the thinning path:
//
//   -md=vol2mesh <in> <out_base> [-onlyBndryVts=false] [-needEuler]
//                [-write2node]                     <out_base>.off: boundary
//                                                  points (default) or mesh
//   -md=vol2skel <in> <out_prefix> [-tt=a,b,c] [-outToDotma]
//                                                  EDT + measure + thinning;
//                                                  per threshold t:
//                                                  <prefix>_thinned<t>.ply
//                                                  and .r (and .ma)
//   -md=vol2ma <in> <out.ply|out.off> [-tt=a,b,c] [-fullOrPruned=0|1|2]
//              [-outToDotma] [-needEuler]          voxel core: the discrete
//                                                  Voronoi diagram of the
//                                                  feature transform (in
//                                                  place of TetGen's), tagged,
//                                                  merged, its inside part
//                                                  FIFO-pruned per threshold;
//                                                  the diagram stays on the
//                                                  device (-hostVoro: copied
//                                                  to host arrays between steps)
//   -md=topo -meshToCheck=<m.off>                  closed pockets of a mesh
//
// Volumes: .mrc (modes 0, 1, 2, 6), .sof / .sog octrees, or raw uint8 with
// -dims=nx,ny,nz (C order [x][y][z]).  Occupancy = value > 0.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/vc_export.h"
#include "exporters.h"
#include "highlevelalgo.h"
#include "octree.h"
#include "surfacing.h"

namespace
{

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

std::vector<double> split_doubles(const std::string& s)
{
    std::vector<double> out;
    std::stringstream ss(s);
    std::string tok;
    while (std::getline(ss, tok, ','))
        if (!tok.empty())
            out.push_back(std::stod(tok));
    return out;
}

bool ends_with(const std::string& s, const std::string& suf)
{
    return s.size() >= suf.size() &&
           s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
}

// MRC: 1024-byte header (nx, ny, nz, mode, ..., nsymbt at byte 92), data with
// x fastest; transposed into C order [x][y][z]
std::shared_ptr<DenseVolume> read_mrc(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f)
        return nullptr;
    int32_t h[256];
    f.read(reinterpret_cast<char*>(h), 1024);
    if (!f)
        return nullptr;
    const int nx = h[0], ny = h[1], nz = h[2], mode = h[3], next = h[23];
    if (nx <= 0 || ny <= 0 || nz <= 0)
        return nullptr;
    f.seekg(1024 + next);
    const size_t n = (size_t)nx * ny * nz;
    std::vector<float> file_order(n);
    for (size_t i = 0; i < n; ++i)
    {
        float v = 0.0f;
        if (mode == 0)
        {
            int8_t b;
            f.read(reinterpret_cast<char*>(&b), 1);
            v = b;
        }
        else if (mode == 1)
        {
            int16_t b;
            f.read(reinterpret_cast<char*>(&b), 2);
            v = b;
        }
        else if (mode == 2)
            f.read(reinterpret_cast<char*>(&v), 4);
        else if (mode == 6)
        {
            uint16_t b;
            f.read(reinterpret_cast<char*>(&b), 2);
            v = b;
        }
        else
            return nullptr;
        file_order[i] = v;
    }
    if (!f)
        return nullptr;
    std::vector<float> c_order(n);
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
                c_order[((size_t)x * ny + y) * nz + z] =
                    file_order[((size_t)z * ny + y) * nx + x];
    return std::make_shared<DenseVolume>(nx, ny, nz, std::move(c_order));
}

std::shared_ptr<DenseVolume> read_raw(const std::string& path, const int d[3])
{
    std::ifstream f(path, std::ios::binary);
    if (!f || d[0] <= 0 || d[1] <= 0 || d[2] <= 0)
        return nullptr;
    std::vector<uint8_t> b((size_t)d[0] * d[1] * d[2]);
    f.read(reinterpret_cast<char*>(b.data()), (std::streamsize)b.size());
    if (!f)
        return nullptr;
    return DenseVolume::fromOccupancy(d[0], d[1], d[2], b.data());
}

bool read_off(const std::string& path, std::vector<point>& v,
              std::vector<std::vector<int>>& faces)
{
    std::ifstream f(path);
    std::string magic;
    size_t nv = 0, nf = 0, ne = 0;
    if (!(f >> magic) || magic != "OFF" || !(f >> nv >> nf >> ne))
        return false;
    v.resize(nv);
    for (auto& p : v)
        f >> p[0] >> p[1] >> p[2];
    faces.resize(nf);
    for (auto& fc : faces)
    {
        int k = 0;
        f >> k;
        fc.resize(k);
        for (auto& i : fc)
            f >> i;
    }
    return (bool)f;
}

bool flag_on(const std::map<std::string, std::string>& flags, const std::string& name,
             bool dflt)
{
    if (flags.count("no" + name))
        return false;
    auto it = flags.find(name);
    if (it == flags.end())
        return dflt;
    return !(it->second == "false" || it->second == "0");
}

int usage()
{
    std::cerr << "usage: voroUtility -md=vol2mesh <in.mrc|in.raw|in.sof|in.sog> "
                 "<out_base> [-onlyBndryVts=false] [-needEuler] [-write2node] "
                 "[-dims=nx,ny,nz]\n"
                 "       voroUtility -md=vol2skel <in> <out_prefix> "
                 "[-tt=0.01,0.02,0.04] [-outToDotma] [-dims=...]\n"
                 "       voroUtility -md=vol2ma <in> <out.ply|out.off> "
                 "[-tt=0.04] [-fullOrPruned=0|1|2] [-outToDotma] [-needEuler] "
                 "[-hostVoro]\n"
                 "       voroUtility -md=topo -meshToCheck=<mesh.off>\n";
    return 1;
}

} // namespace

int main(int argc, char** argv)
{
    std::map<std::string, std::string> flags;
    std::vector<std::string> pos;
    for (int i = 1; i < argc; ++i)
    {
        std::string a = argv[i];
        if (a.size() > 1 && a[0] == '-')
        {
            a = a.substr(a[1] == '-' ? 2 : 1);
            const size_t eq = a.find('=');
            if (eq == std::string::npos)
                flags[a] = "true";
            else
                flags[a.substr(0, eq)] = a.substr(eq + 1);
        }
        else
            pos.push_back(a);
    }
    const std::string md = flags.count("md") ? flags["md"] : "";
    try
    {
        auto load = [&](const std::string& path) -> std::shared_ptr<Volume3DScalar> {
            if (ends_with(path, ".mrc"))
                return read_mrc(path);
            if (ends_with(path, ".sof") || ends_with(path, ".sog"))
                return std::make_shared<OctreeVolume>(path.c_str());
            int d[3] = {0, 0, 0};
            if (flags.count("dims"))
            {
                const auto v = split_doubles(flags["dims"]);
                for (int k = 0; k < 3 && k < (int)v.size(); ++k)
                    d[k] = (int)v[k];
            }
            return read_raw(path, d);
        };
        if (md == "vol2mesh")
        {
            if (pos.size() < 2)
                return usage();
            auto vol = load(pos[0]);
            if (!vol)
            {
                std::cerr << "cannot read volume " << pos[0] << "\n";
                return 1;
            }
            const auto t0 = Clock::now();
            const bool need_euler = flag_on(flags, "needEuler", false);
            const bool node = flag_on(flags, "write2node", false);
            const auto rc =
                flag_on(flags, "onlyBndryVts", true)
                    ? voxelvoro::writeVolumeAsBoundaryPts(vol, pos[1].c_str(), node)
                    : voxelvoro::writeVolumeAsBoundaryMesh(vol, pos[1].c_str(),
                                                           need_euler, node);
            std::cout << "time -> extract + write boundary: " << ms_since(t0) << " ms\n";
            if (rc != voxelvoro::ExportErrCode::SUCCESS)
            {
                std::cerr << "failed to write boundary mesh from " << pos[0] << " to "
                          << pos[1] << "\n";
                return 1;
            }
            return 0;
        }
        if (md == "vol2skel")
        {
            if (pos.size() < 2)
                return usage();
            auto vol = load(pos[0]);
            if (!vol)
            {
                std::cerr << "cannot read volume " << pos[0] << "\n";
                return 1;
            }
            std::vector<double> tt = {0.04};
            if (flags.count("tt"))
                tt = split_doubles(flags["tt"]);
            const auto t0 = Clock::now();
            VoxelThinning thin;
            std::vector<float> levels;
            voxelvoro::thinVolume(vol, tt, thin, &levels);
            std::cout << "time(THIN) -> EDT + measure + thinning: " << ms_since(t0)
                      << " ms\n";
            std::sort(tt.begin(), tt.end());
            const xform& mat = vol->getVoxToModelMat();
            const std::vector<int32_t> d2 = thin.squaredDistances();
            const bool dotma = flag_on(flags, "outToDotma", false);
            const int ny = thin.sizeY(), nz = thin.sizeZ();
            for (size_t k = 0; k < levels.size(); ++k)
            {
                std::vector<ivec3> vox;
                thin.getRemainedVoxels((int)k, vox);
                std::ostringstream name;
                name << pos[1] << "_thinned" << tt[k];
                std::cout << "level " << levels[k] << " (t=" << tt[k]
                          << "): " << vox.size() << " voxels -> " << name.str() << ".ply\n";
                // skeleton voxels in model space; radius = distance to the
                // nearest outside voxel, sqrt of the exact squared EDT
                std::vector<point> v, v_model;
                std::vector<float> r, r_model;
                const point o = mat * point(0.0f, 0.0f, 0.0f);
                for (const auto& p : vox)
                {
                    v.push_back(point((float)p[0], (float)p[1], (float)p[2]));
                    v_model.push_back(mat * v.back());
                    const size_t i = ((size_t)p[0] * ny + p[1]) * nz + p[2];
                    r.push_back(std::sqrt((float)d2[i]));
                    const point rp = mat * point(r.back(), 0.0f, 0.0f);
                    const float d0 = rp[0] - o[0], d1 = rp[1] - o[1], dz = rp[2] - o[2];
                    r_model.push_back(std::sqrt(d0 * d0 + d1 * d1 + dz * dz));
                }
                if (voxelvoro::writeToPLY((name.str() + ".ply").c_str(), v_model, {}, {}, {},
                                          {}, {}) != voxelvoro::ExportErrCode::SUCCESS)
                    return 1;
                std::vector<float> flat(3 * v.size() + 3);
                for (size_t i = 0; i < v.size(); ++i)
                    for (int a = 0; a < 3; ++a)
                        flat[3 * i + a] = v[i][a];
                if (vc_write_radii((name.str() + ".r").c_str(), flat.data(), r.data(),
                                   (int64_t)v.size(), mat.data()) != VC_OK)
                    return 1;
                if (dotma && voxelvoro::writeToDotma(name.str() + ".ma", v_model, {}, {},
                                                     r_model) !=
                                 voxelvoro::ExportErrCode::SUCCESS)
                    return 1;
            }
            return 0;
        }
        if (md == "vol2ma")
        {
            // src/voroUtility.cpp:454-525
            if (pos.size() < 2)
                return usage();
            auto vol = load(pos[0]);
            if (!vol)
            {
                std::cerr << "cannot read volume " << pos[0] << "\n";
                return 1;
            }
            const bool need_euler = flag_on(flags, "needEuler", false);
            std::vector<float> tt = {0.04f};
            if (flags.count("tt"))
            {
                // the reference parses each threshold with std::stof
                tt.clear();
                std::stringstream ss(flags["tt"]);
                std::string tok;
                while (std::getline(ss, tok, ','))
                    tt.push_back(std::stof(tok));
            }
            const int full_or_pruned =
                flags.count("fullOrPruned") ? std::stoi(flags["fullOrPruned"]) : 0;
            const auto t0 = Clock::now();
            voxelvoro::VoroInfo voro;
            // -hostVoro: the diagram on host arrays between the steps
            if (flag_on(flags, "hostVoro", false))
                voxelvoro::computeVDHost(vol, voro);
            else
                voxelvoro::computeVD(vol, voro);
            std::cout << "time -> discrete voronoi diagram: " << ms_since(t0)
                      << " ms\n";
            if (!voxelvoro::preprocessVoro(voro, vol, need_euler))
            {
                std::cerr << "Error processing voro!\n";
                return 1;
            }
            const auto t1 = Clock::now();
            if (voxelvoro::exportInsideVoroMesh(
                    voro, vol, pos[1].c_str(), need_euler,
                    flag_on(flags, "collapseZeroLenEdges", false), true, false,
                    full_or_pruned, true, flag_on(flags, "outToDotma", false),
                    tt) != voxelvoro::ExportErrCode::SUCCESS)
            {
                std::cerr << "Error: couldn't write inside voro info.\n";
                return 1;
            }
            std::cout << "time -> extract + prune + write: " << ms_since(t1)
                      << " ms\n";
            return 0;
        }
        if (md == "topo")
        {
            if (!flags.count("meshToCheck"))
                return usage();
            std::vector<point> v;
            std::vector<std::vector<int>> faces;
            if (!read_off(flags["meshToCheck"], v, faces))
            {
                std::cerr << "cannot read mesh\n";
                return 1;
            }
            cellcomplex cc(v, {}, {}, false);
            for (const auto& f : faces)
                cc.appendFace(f);
            cc.finalize();
            eulerchar ec;
            cc.eulerChar(ec);
            ec.logToConsole("input mesh");
            std::cout << "# closed components: " << voxelvoro::nClosedComponents(cc)
                      << "\n";
            return 0;
        }
    }
    catch (const std::exception& e)
    {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return usage();
}
