// voroinfo.cpp -- VoroInfo over include/vc_voro.h (see voroinfo.h).
#include "voroinfo.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "../../include/vc_cc.h"
#include "../../include/vc_vdiag.h"
#include "../../include/vc_voro.h"
#include "vc_device.h"

#pragma clang fp contract(off)

float MeasureForMA::lambdaForFace(const point& a, const point& b)
{
    float q = (b[0] - a[0]) * (b[0] - a[0]);
    q += (b[1] - a[1]) * (b[1] - a[1]);
    q += (b[2] - a[2]) * (b[2] - a[2]);
    return std::sqrt(q);
}

namespace voxelvoro
{

void VoroInfo::setGeometry(const std::vector<point>& vts,
                           const std::vector<std::vector<int>>& faces,
                           const std::vector<ivec2>& face_sites,
                           const std::vector<point>& sites,
                           const std::vector<bool>& finite_v)
{
    if (face_sites.size() != faces.size())
        throw std::runtime_error("VoroInfo: one site pair per face");
    m_dev.reset();
    drop_device_copies();
    m_geom = cellcomplex(vts, {}, {}, false);
    for (const auto& f : faces)
        m_geom.appendFace(f);
    m_geom.finalize();
    m_face_sites = face_sites;
    m_site_positions = sites;
    m_is_finite_v = finite_v.empty() ? std::vector<bool>(vts.size(), true) : finite_v;
    m_vts_valid.assign(vts.size(), false);
    m_tagged = false;
    m_only_inside = -1;
    m_r_per_v.clear();
    m_r_valid = false;
    recompute_flags();
}

void VoroInfo::setDevice(std::shared_ptr<vc_vdiag> d)
{
    m_dev = std::move(d);
    drop_device_copies();
    m_geom = cellcomplex();
    m_face_sites.clear();
    m_site_positions.clear();
    m_vts_valid.clear();
    m_is_finite_v.clear();
    m_edge_valid.clear();
    m_face_valid.clear();
    m_tagged = false;
    m_only_inside = -1;
    m_bb_valid = false;
    m_r_per_v.clear();
    m_r_valid = false;
}

namespace
{
vc_vdiag_info dev_info(vc_vdiag* d)
{
    vc_vdiag_info in;
    vcdev::check(vc_vdiag_info_get(d, &in), "vc_vdiag_info_get");
    return in;
}
} // namespace

void VoroInfo::drop_device_copies() const
{
    m_dev_vts.clear();
    m_got_vts = m_got_fs = m_got_sites = m_got_radii = false;
    m_inside_cc.reset();
}

void VoroInfo::own_device()
{
    // another VoroInfo shares the resident diagram: keep it unchanged for
    // that one and continue on the host
    if (m_dev && m_dev.use_count() > 1)
        materialize();
}

void VoroInfo::materialize()
{
    if (!m_dev)
        return;
    vc_ctx* ctx = vcdev::context();
    const vc_vdiag_info in = dev_info(m_dev.get());
    std::vector<float> vts(3 * (size_t)in.nv + 3), sites(3 * (size_t)in.ns + 3),
        radii((size_t)in.nv + 1);
    std::vector<int32_t> edges(2 * (size_t)in.ne + 2), fv((size_t)in.n_sides + 1),
        fs(2 * (size_t)in.nf + 2);
    std::vector<int64_t> fo((size_t)in.nf + 1);
    std::vector<uint8_t> valid((size_t)in.nv + 1);
    vcdev::check(vc_vdiag_get(ctx, m_dev.get(), vts.data(), edges.data(), fo.data(),
                              fv.data(), fs.data(), sites.data(),
                              in.radii ? radii.data() : nullptr, valid.data()),
                 "vc_vdiag_get");
    std::vector<point> pv((size_t)in.nv);
    for (int64_t i = 0; i < in.nv; ++i)
        pv[i] = point(vts[3 * i], vts[3 * i + 1], vts[3 * i + 2]);
    std::vector<ivec2> pe((size_t)in.ne);
    for (int64_t e = 0; e < in.ne; ++e)
        pe[e] = ivec2(edges[2 * e], edges[2 * e + 1]);
    // the finalized edges come first, so finalizing again keeps the numbering
    cellcomplex g(pv, pe, {}, false);
    for (int64_t f = 0; f < in.nf; ++f)
        g.appendFace(std::vector<int>(fv.begin() + fo[f], fv.begin() + fo[f + 1]));
    g.finalize();
    m_geom = std::move(g);
    m_face_sites.resize((size_t)in.nf);
    for (int64_t f = 0; f < in.nf; ++f)
        m_face_sites[f] = ivec2(fs[2 * f], fs[2 * f + 1]);
    m_site_positions.resize((size_t)in.ns);
    for (int64_t i = 0; i < in.ns; ++i)
        m_site_positions[i] = point(sites[3 * i], sites[3 * i + 1], sites[3 * i + 2]);
    m_is_finite_v.assign((size_t)in.nv, true);
    m_vts_valid.resize((size_t)in.nv);
    for (int64_t i = 0; i < in.nv; ++i)
        m_vts_valid[i] = valid[i] != 0;
    if (in.radii)
        m_r_per_v.assign(radii.begin(), radii.begin() + in.nv);
    m_dev.reset();
    drop_device_copies();
    recompute_flags();
}

size_t VoroInfo::numVts() const
{
    return m_dev ? (size_t)dev_info(m_dev.get()).nv : m_geom.numVts();
}
size_t VoroInfo::numEdges() const
{
    return m_dev ? (size_t)dev_info(m_dev.get()).ne : m_geom.numEdges();
}
size_t VoroInfo::numFaces() const
{
    return m_dev ? (size_t)dev_info(m_dev.get()).nf : m_geom.numFaces();
}

const point& VoroInfo::vertex(int vi) const
{
    if (!m_dev)
        return m_geom.getVert(vi);
    if (!m_got_vts)
    {
        const vc_vdiag_info in = dev_info(m_dev.get());
        std::vector<float> v(3 * (size_t)in.nv + 3);
        vcdev::check(vc_vdiag_get(vcdev::context(), m_dev.get(), v.data(), nullptr, nullptr,
                                  nullptr, nullptr, nullptr, nullptr, nullptr),
                     "vc_vdiag_get");
        m_dev_vts.resize((size_t)in.nv);
        for (int64_t i = 0; i < in.nv; ++i)
            m_dev_vts[i] = point(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
        m_got_vts = true;
    }
    return m_dev_vts[vi];
}

const std::vector<ivec2>& VoroInfo::faceSites() const
{
    if (m_dev && !m_got_fs)
    {
        const vc_vdiag_info in = dev_info(m_dev.get());
        std::vector<int32_t> fs(2 * (size_t)in.nf + 2);
        vcdev::check(vc_vdiag_get(vcdev::context(), m_dev.get(), nullptr, nullptr, nullptr,
                                  nullptr, fs.data(), nullptr, nullptr, nullptr),
                     "vc_vdiag_get");
        auto& out = const_cast<std::vector<ivec2>&>(m_face_sites);
        out.resize((size_t)in.nf);
        for (int64_t f = 0; f < in.nf; ++f)
            out[f] = ivec2(fs[2 * f], fs[2 * f + 1]);
        m_got_fs = true;
    }
    return m_face_sites;
}

const std::vector<point>& VoroInfo::getSitesPosition() const
{
    if (m_dev && !m_got_sites)
    {
        const vc_vdiag_info in = dev_info(m_dev.get());
        std::vector<float> s(3 * (size_t)in.ns + 3);
        vcdev::check(vc_vdiag_get(vcdev::context(), m_dev.get(), nullptr, nullptr, nullptr,
                                  nullptr, nullptr, s.data(), nullptr, nullptr),
                     "vc_vdiag_get");
        auto& out = const_cast<std::vector<point>&>(m_site_positions);
        out.resize((size_t)in.ns);
        for (int64_t i = 0; i < in.ns; ++i)
            out[i] = point(s[3 * i], s[3 * i + 1], s[3 * i + 2]);
        m_got_sites = true;
    }
    return m_site_positions;
}

const std::vector<float>& VoroInfo::getRadii() const
{
    if (m_dev && m_r_valid && !m_got_radii)
    {
        const vc_vdiag_info in = dev_info(m_dev.get());
        auto& out = const_cast<std::vector<float>&>(m_r_per_v);
        out.assign((size_t)in.nv + 1, 0.0f);
        vcdev::check(vc_vdiag_get(vcdev::context(), m_dev.get(), nullptr, nullptr, nullptr,
                                  nullptr, nullptr, nullptr, out.data(), nullptr),
                     "vc_vdiag_get");
        out.resize((size_t)in.nv);
        m_got_radii = true;
    }
    return m_r_per_v;
}

std::shared_ptr<vc_cc> VoroInfo::takeInsideComplex() const
{
    std::shared_ptr<vc_cc> r;
    r.swap(m_inside_cc);
    return r;
}

namespace
{
// face polygons as CSR (offsets, vertex ids), one spare entry at the end
void face_csr(const cellcomplex& g, std::vector<int64_t>& foff,
              std::vector<int32_t>& fvts)
{
    foff.assign(1, 0);
    fvts.clear();
    std::vector<int> fv;
    for (size_t f = 0; f < g.numFaces(); ++f)
    {
        g.getFaceVRep((int)f, fv);
        fvts.insert(fvts.end(), fv.begin(), fv.end());
        foff.push_back((int64_t)fvts.size());
    }
    fvts.push_back(0);
}
} // namespace

void VoroInfo::computeRadii()
{
    if (m_dev)
    {
        own_device();
    }
    if (m_dev)
    {
        vcdev::check(vc_vdiag_radii(vcdev::context(), m_dev.get()), "vc_vdiag_radii");
        m_r_valid = true;
        m_got_radii = false;
        return;
    }
    const int nv = (int)m_geom.numVts(), nf = (int)m_geom.numFaces();
    std::vector<float> vts(3 * (size_t)nv + 3), sites(3 * m_site_positions.size() + 3);
    for (int v = 0; v < nv; ++v)
        for (int a = 0; a < 3; ++a)
            vts[3 * v + a] = m_geom.getVert(v)[a];
    for (size_t i = 0; i < m_site_positions.size(); ++i)
        for (int a = 0; a < 3; ++a)
            sites[3 * i + a] = m_site_positions[i][a];
    std::vector<int64_t> foff;
    std::vector<int32_t> fvts, fs(2 * (size_t)nf + 2);
    face_csr(m_geom, foff, fvts);
    for (int f = 0; f < nf; ++f)
    {
        fs[2 * f] = m_face_sites[f][0];
        fs[2 * f + 1] = m_face_sites[f][1];
    }
    m_r_per_v.assign((size_t)nv + 1, 0.0f);
    vcdev::check(vc_voro_radii(vcdev::context(), nv, nf, vts.data(), foff.data(),
                               fvts.data(), sites.data(),
                               (int32_t)m_site_positions.size(), fs.data(),
                               m_r_per_v.data()),
                 "vc_voro_radii");
    m_r_per_v.resize((size_t)nv);
    m_r_valid = true;
}

void VoroInfo::getPartInside(std::vector<int>& vts_indices,
                             std::vector<ivec2>& edges,
                             std::vector<uTriFace>& tri_faces) const
{
    ensure_host();
    const int nv = (int)m_geom.numVts(), nf = (int)m_geom.numFaces();
    std::vector<int64_t> foff;
    std::vector<int32_t> fvts;
    face_csr(m_geom, foff, fvts);
    std::vector<uint8_t> valid((size_t)nv + 1, 0);
    for (int v = 0; v < nv; ++v)
        valid[v] = isVertexValid(v) ? 1 : 0;
    int64_t n1 = 0, n2 = 0, n3 = 0;
    int rc = vc_voro_part_inside(nv, nf, foff.data(), fvts.data(), valid.data(),
                                 nullptr, 0, &n1, nullptr, 0, &n2, nullptr, 0, &n3);
    if (rc != VC_OK && rc != VC_ERR_CAPACITY)
        vcdev::check(rc, "vc_voro_part_inside");
    std::vector<int32_t> vi(n1 + 1), eo(2 * n2 + 2);
    std::vector<uint32_t> tr(3 * n3 + 3);
    vcdev::check(vc_voro_part_inside(nv, nf, foff.data(), fvts.data(), valid.data(),
                                     vi.data(), n1, &n1, eo.data(), n2, &n2,
                                     tr.data(), n3, &n3),
                 "vc_voro_part_inside");
    for (int64_t i = 0; i < n1; ++i)
        vts_indices.push_back(vi[i]);
    for (int64_t i = 0; i < n2; ++i)
        edges.push_back(ivec2(eo[2 * i], eo[2 * i + 1]));
    for (int64_t i = 0; i < n3; ++i)
        tri_faces.push_back(uTriFace(tr[3 * i], tr[3 * i + 1], tr[3 * i + 2]));
}

bool VoroInfo::isVertexValid(int vi) const
{
    ensure_host();
    return onlyHasInside() || (m_vts_valid[vi] && m_is_finite_v[vi]);
}
bool VoroInfo::isEdgeValid(int ei) const
{
    ensure_host();
    return onlyHasInside() || m_edge_valid[ei];
}
bool VoroInfo::isFaceValid(int fi) const
{
    ensure_host();
    return onlyHasInside() || m_face_valid[fi];
}

void VoroInfo::recompute_flags()
{
    m_edge_valid.assign(m_geom.numEdges(), false);
    for (size_t e = 0; e < m_geom.numEdges(); ++e)
    {
        const ivec2& ed = m_geom.getEdge(e);
        m_edge_valid[e] = isVertexValid(ed[0]) && isVertexValid(ed[1]);
    }
    m_face_valid.assign(m_geom.numFaces(), false);
    std::vector<int> fv;
    for (size_t f = 0; f < m_geom.numFaces(); ++f)
    {
        m_geom.getFaceVRep((int)f, fv);
        bool ok = true;
        for (int v : fv)
            ok = ok && isVertexValid(v);
        m_face_valid[f] = ok;
    }
}

bool VoroInfo::tagVert(const point& p,
                       const std::shared_ptr<Volume3DScalar>& vol) const
{
    const int v[3] = {(int)std::round(p[0]), (int)std::round(p[1]),
                      (int)std::round(p[2])};
    if (v[0] < 0 || v[1] < 0 || v[2] < 0 || v[0] >= vol->getSizeX() ||
        v[1] >= vol->getSizeY() || v[2] >= vol->getSizeZ())
        return false;
    return vol->getDataAt(v[0], v[1], v[2]) > 0.0;
}

bool VoroInfo::tagVtsUsingUniformVol(const std::shared_ptr<Volume3DScalar>& vol)
{
    vc_ctx* ctx = vcdev::context();
    const std::vector<uint8_t> occ = vol->occupancy();
    vc_vol* dv = nullptr;
    vcdev::check(vc_volume_upload(ctx, occ.data(), vol->getSizeX(),
                                  vol->getSizeY(), vol->getSizeZ(), &dv),
                 "vc_volume_upload");
    own_device();
    if (m_dev)
    {
        const int rc = vc_vdiag_tag(ctx, m_dev.get(), dv);
        vc_volume_release(ctx, dv);
        vcdev::check(rc, "vc_vdiag_tag");
        m_tagged = true;
        const vc_vdiag_info in = dev_info(m_dev.get());
        // the valid vertices' box (every vertex is finite here)
        m_bb_valid = onlyHasInside() ? in.nv > 0 : in.n_valid > 0;
        if (onlyHasInside() && in.nv > 0 && in.n_valid != in.nv)
        {
            // every vertex counts as valid: the box of all of them
            materialize();
            return tagVtsUsingUniformVol(vol);
        }
        for (int a = 0; a < 3; ++a)
        {
            m_bb_min[a] = in.bb_min[a];
            m_bb_max[a] = in.bb_max[a];
        }
        return true;
    }
    const size_t n = m_geom.numVts();
    std::vector<float> pts(3 * n);
    for (size_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a)
            pts[3 * i + a] = m_geom.getVert(i)[a];
    std::vector<uint8_t> valid(n + 1, 0);
    const int rc = vc_voro_tag_points(ctx, dv, pts.data(), (int64_t)n, valid.data());
    vc_volume_release(ctx, dv);
    vcdev::check(rc, "vc_voro_tag_points");
    for (size_t i = 0; i < n; ++i)
        m_vts_valid[i] = valid[i] != 0;
    m_tagged = true;
    recompute_flags();
    // bounding box of the valid vertices
    m_bb_valid = false;
    for (size_t i = 0; i < n; ++i)
    {
        if (!isVertexValid((int)i))
            continue;
        const point& p = m_geom.getVert(i);
        if (!m_bb_valid)
        {
            m_bb_min = m_bb_max = p;
            m_bb_valid = true;
        }
        for (int a = 0; a < 3; ++a)
        {
            m_bb_min[a] = std::min(m_bb_min[a], p[a]);
            m_bb_max[a] = std::max(m_bb_max[a], p[a]);
        }
    }
    return true;
}

bool VoroInfo::isFaceFinite(int fi) const
{
    // a resident diagram is the discrete one, whose vertices are all finite
    if (onlyHasInside() || m_dev)
        return true;
    std::vector<int> poly;
    m_geom.getFaceVRep(fi, poly);
    for (int v : poly)
        if (!m_is_finite_v[v])
            return false;
    return true;
}

float VoroInfo::getInsidePartSize() const
{
    if (!m_bb_valid)
        return 0.0f;
    float s = 0.0f;
    for (int a = 0; a < 3; ++a)
        s = std::max(s, m_bb_max[a] - m_bb_min[a]);
    return s;
}

void VoroInfo::getValidVts(std::vector<int>& idx) const
{
    idx.clear();
    if (m_dev)
    {
        const vc_vdiag_info in = dev_info(m_dev.get());
        std::vector<uint8_t> valid((size_t)in.nv + 1);
        vcdev::check(vc_vdiag_get(vcdev::context(), m_dev.get(), nullptr, nullptr, nullptr,
                                  nullptr, nullptr, nullptr, nullptr, valid.data()),
                     "vc_vdiag_get");
        for (int64_t v = 0; v < in.nv; ++v)
            if (onlyHasInside() || valid[v])
                idx.push_back((int)v);
        return;
    }
    for (size_t v = 0; v < m_geom.numVts(); ++v)
        if (isVertexValid((int)v))
            idx.push_back((int)v);
}
void VoroInfo::getValidEdges(std::vector<int>& idx) const
{
    ensure_host();
    idx.clear();
    for (size_t e = 0; e < m_geom.numEdges(); ++e)
        if (isEdgeValid((int)e))
            idx.push_back((int)e);
}
void VoroInfo::getValidFaces(std::vector<int>& idx) const
{
    ensure_host();
    idx.clear();
    for (size_t f = 0; f < m_geom.numFaces(); ++f)
        if (isFaceValid((int)f))
            idx.push_back((int)f);
}

void VoroInfo::all_measures(std::vector<float>& lv, std::vector<float>& le,
                            std::vector<float>& lf) const
{
    ensure_host();
    const int nv = (int)m_geom.numVts(), ne = (int)m_geom.numEdges(),
              nf = (int)m_geom.numFaces();
    std::vector<int64_t> ve_off(nv + 1, 0), ef_off(ne + 1, 0);
    std::vector<int32_t> ve, ef;
    for (int v = 0; v < nv; ++v)
    {
        for (int k = 0; k < m_geom.cntNbEdgesofVert(v); ++k)
            ve.push_back(m_geom.nbEdgeofVert(v, k));
        ve_off[v + 1] = (int64_t)ve.size();
    }
    for (int e = 0; e < ne; ++e)
    {
        for (int k = 0; k < m_geom.cntNbFacesofEdge(e); ++k)
            ef.push_back(m_geom.nbFaceofEdge(e, k));
        ef_off[e + 1] = (int64_t)ef.size();
    }
    std::vector<float> sites(3 * m_site_positions.size() + 3);
    for (size_t i = 0; i < m_site_positions.size(); ++i)
        for (int a = 0; a < 3; ++a)
            sites[3 * i + a] = m_site_positions[i][a];
    std::vector<int32_t> fs(2 * (size_t)nf + 2);
    std::vector<uint8_t> fv((size_t)nf + 1);
    for (int f = 0; f < nf; ++f)
    {
        fs[2 * f] = m_face_sites[f][0];
        fs[2 * f + 1] = m_face_sites[f][1];
        fv[f] = isFaceValid(f) ? 1 : 0;
    }
    ve.push_back(0);
    ef.push_back(0);
    lv.assign((size_t)nv + 1, 0.0f);
    le.assign((size_t)ne + 1, 0.0f);
    lf.assign((size_t)nf + 1, 0.0f);
    vcdev::check(vc_voro_measures(vcdev::context(), nv, ne, nf, sites.data(),
                                  (int32_t)m_site_positions.size(), fs.data(),
                                  fv.data(), ve_off.data(), ve.data(),
                                  ef_off.data(), ef.data(), lv.data(),
                                  le.data(), lf.data()),
                 "vc_voro_measures");
}

void VoroInfo::computeVertexMeasure(MeasureForMA::meassuretype tp,
                                    const std::vector<int>& idx,
                                    std::vector<float>& out) const
{
    out.clear();
    if (tp != MeasureForMA::LAMBDA)
        return;
    std::vector<float> lv, le, lf;
    all_measures(lv, le, lf);
    for (int v : idx)
        out.push_back(lv[v]);
}
void VoroInfo::computeEdgesMeasure(MeasureForMA::meassuretype tp,
                                   const std::vector<int>& idx,
                                   std::vector<float>& out) const
{
    out.clear();
    if (tp != MeasureForMA::LAMBDA)
        return;
    std::vector<float> lv, le, lf;
    all_measures(lv, le, lf);
    for (int e : idx)
        out.push_back(le[e]);
}
void VoroInfo::computeFacesMeasure(MeasureForMA::meassuretype tp,
                                   const std::vector<int>& idx,
                                   std::vector<float>& out) const
{
    out.clear();
    if (tp != MeasureForMA::LAMBDA)
        return;
    std::vector<float> lv, le, lf;
    all_measures(lv, le, lf);
    for (int f : idx)
        out.push_back(lf[f]);
}

void VoroInfo::extractInsideWithMeasure(
    MeasureForMA::meassuretype tp, std::vector<point>& out_vts,
    std::vector<ivec2>& out_edges, std::vector<uTriFace>& out_tris,
    std::vector<int>& from_fi, std::vector<float>& vts_msure,
    std::vector<float>& edges_msure, std::vector<float>& faces_msure) const
{
    out_vts.clear();
    out_edges.clear();
    out_tris.clear();
    from_fi.clear();
    vts_msure.clear();
    edges_msure.clear();
    faces_msure.clear();
    if (tp != MeasureForMA::LAMBDA)
        return;
    if (m_dev)
    {
        // on the resident diagram; the output's finalized complex is kept
        // for the collapse (takeInsideComplex)
        const vc_vdiag_info in = dev_info(m_dev.get());
        // faces have >= 3 vertices: sum(size - 2) triangles
        const int64_t tcap = std::max<int64_t>(in.n_sides - 2 * in.nf, 0);
        const int64_t vcap = in.nv, ecap = in.ne + 3 * tcap;
        std::unique_ptr<int32_t[]> vi(new int32_t[vcap + 1]), eo(new int32_t[2 * ecap + 2]),
            tr(new int32_t[3 * tcap + 3]), ff(new int32_t[tcap + 1]);
        std::unique_ptr<float[]> vm(new float[vcap + 1]), em(new float[ecap + 1]),
            fm(new float[tcap + 1]);
        int64_t n1 = 0, n2 = 0, n3 = 0;
        vc_cc* inside = nullptr;
        vc_ctx* ctx = vcdev::context();
        vcdev::check(vc_vdiag_extract_inside(ctx, m_dev.get(), onlyHasInside() ? 1 : 0,
                                             vi.get(), vm.get(), vcap, &n1, eo.get(),
                                             em.get(), ecap, &n2, tr.get(), ff.get(),
                                             fm.get(), tcap, &n3, &inside),
                     "vc_vdiag_extract_inside");
        m_inside_cc = std::shared_ptr<vc_cc>(inside, [ctx](vc_cc* p) { vc_cc_release(ctx, p); });
        out_vts.reserve(n1);
        vts_msure.assign(vm.get(), vm.get() + n1);
        for (int64_t i = 0; i < n1; ++i)
            out_vts.push_back(vertex(vi[i]));
        out_edges.reserve(n2);
        for (int64_t i = 0; i < n2; ++i)
            out_edges.push_back(ivec2(eo[2 * i], eo[2 * i + 1]));
        edges_msure.assign(em.get(), em.get() + n2);
        out_tris.reserve(n3);
        for (int64_t i = 0; i < n3; ++i)
            out_tris.push_back(uTriFace((unsigned)tr[3 * i], (unsigned)tr[3 * i + 1],
                                        (unsigned)tr[3 * i + 2]));
        from_fi.assign(ff.get(), ff.get() + n3);
        faces_msure.assign(fm.get(), fm.get() + n3);
        return;
    }
    std::vector<float> lv, le, lf;
    all_measures(lv, le, lf);
    const int nv = (int)m_geom.numVts(), ne = (int)m_geom.numEdges(),
              nf = (int)m_geom.numFaces();
    const std::vector<int32_t> edges = m_geom.edgeArray();
    std::vector<int64_t> foff(1, 0);
    std::vector<int32_t> fvts, fv;
    for (int f = 0; f < nf; ++f)
    {
        m_geom.getFaceVRep(f, fv);
        fvts.insert(fvts.end(), fv.begin(), fv.end());
        foff.push_back((int64_t)fvts.size());
    }
    fvts.push_back(0);
    std::vector<uint8_t> valid((size_t)nv + 1, 0);
    for (int v = 0; v < nv; ++v)
        valid[v] = isVertexValid(v) ? 1 : 0;
    // on the device, with the upper bounds as capacities (one pass): every
    // vertex, every triangle of the zig-zag triangulations, every edge plus
    // three sides per triangle
    int64_t nt_max = 0;
    for (int32_t f = 0; f < nf; ++f)
        nt_max += std::max<int64_t>(foff[f + 1] - foff[f] - 2, 0);
    int64_t n1 = nv, n2 = ne + 3 * nt_max, n3 = nt_max;
    std::vector<int32_t> vi(n1 + 1), eo(2 * n2 + 2), tr(3 * n3 + 3), ff(n3 + 1);
    std::vector<float> em(n2 + 1), fm(n3 + 1);
    vcdev::check(vc_voro_extract_inside_device(
                     vcdev::context(), nv, ne, nf, edges.data(), foff.data(),
                     fvts.data(), valid.data(), le.data(), lf.data(), vi.data(),
                     n1, &n1, eo.data(), em.data(), n2, &n2, tr.data(), ff.data(),
                     fm.data(), n3, &n3),
                 "vc_voro_extract_inside_device");
    for (int64_t i = 0; i < n1; ++i)
    {
        out_vts.push_back(m_geom.getVert(vi[i]));
        vts_msure.push_back(lv[vi[i]]);
    }
    for (int64_t i = 0; i < n2; ++i)
    {
        out_edges.push_back(ivec2(eo[2 * i], eo[2 * i + 1]));
        edges_msure.push_back(em[i]);
    }
    for (int64_t i = 0; i < n3; ++i)
    {
        out_tris.push_back(uTriFace((unsigned)tr[3 * i], (unsigned)tr[3 * i + 1],
                                    (unsigned)tr[3 * i + 2]));
        from_fi.push_back(ff[i]);
        faces_msure.push_back(fm[i]);
    }
}

void VoroInfo::mergeCloseVts(float eps, bool only_inside)
{
    own_device();
    if (m_dev)
    {
        int on_device = 0;
        vcdev::check(vc_vdiag_merge_close(vcdev::context(), m_dev.get(), eps,
                                          only_inside ? 1 : 0, &on_device),
                     "vc_vdiag_merge_close");
        if (on_device)
        {
            // the resident diagram is the merged one: seeds keep their
            // validity and radius (the valid part only, with only_inside)
            drop_device_copies();
            return;
        }
        // vertices merge: the order-defined BFS on the host
        materialize();
    }
    const int nv = (int)m_geom.numVts(), ne = (int)m_geom.numEdges(),
              nf = (int)m_geom.numFaces();
    std::vector<float> vts(3 * (size_t)nv + 3);
    for (int v = 0; v < nv; ++v)
        for (int a = 0; a < 3; ++a)
            vts[3 * v + a] = m_geom.getVert(v)[a];
    const std::vector<int32_t> edges = m_geom.edgeArray();
    std::vector<int64_t> foff(1, 0), ve_off(nv + 1, 0);
    std::vector<int32_t> fvts, fv, ve;
    for (int f = 0; f < nf; ++f)
    {
        m_geom.getFaceVRep(f, fv);
        fvts.insert(fvts.end(), fv.begin(), fv.end());
        foff.push_back((int64_t)fvts.size());
    }
    for (int v = 0; v < nv; ++v)
    {
        for (int k = 0; k < m_geom.cntNbEdgesofVert(v); ++k)
            ve.push_back(m_geom.nbEdgeofVert(v, k));
        ve_off[v + 1] = (int64_t)ve.size();
    }
    fvts.push_back(0);
    ve.push_back(0);
    std::vector<uint8_t> valid((size_t)nv + 1, 0);
    for (int v = 0; v < nv; ++v)
        valid[v] = isVertexValid(v) ? 1 : 0;
    std::vector<int32_t> o2n((size_t)nv + 1), nvts((size_t)nv + 1),
        nfv(fvts.size() + 1), ned(2 * ((size_t)ne + nf) + 2);
    std::vector<int64_t> nfo((size_t)nf + 2);
    std::vector<uint8_t> rem((size_t)nf + 1), keep((size_t)ne + 1);
    int32_t n_new = 0;
    int64_t n_e = 0;
    // on the device when nothing lies within eps along an edge, else the
    // order-defined BFS on the host
    vcdev::check(vc_voro_merge_close_device(
                     vcdev::context(), nv, ne, nf, vts.data(), edges.data(),
                     foff.data(), fvts.data(), ve_off.data(), ve.data(),
                     valid.data(), eps, only_inside ? 1 : 0, o2n.data(),
                     nvts.data(), &n_new, nfo.data(), nfv.data(), rem.data(),
                     ned.data(), (int64_t)ne + nf, &n_e, keep.data(), nullptr),
                 "vc_voro_merge_close_device");
    // the new geometry: seeds, surviving edges, kept faces, then finalize
    std::vector<point> new_v;
    std::vector<bool> new_valid, new_finite;
    for (int32_t i = 0; i < n_new; ++i)
    {
        new_v.push_back(m_geom.getVert(nvts[i]));
        new_valid.push_back(m_vts_valid[nvts[i]]);
        new_finite.push_back(onlyHasInside() || m_is_finite_v[nvts[i]]);
    }
    std::vector<ivec2> new_e;
    for (int64_t i = 0; i < n_e; ++i)
        new_e.push_back(ivec2(ned[2 * i], ned[2 * i + 1]));
    cellcomplex g(new_v, new_e, {}, false);
    int kept = 0;
    std::vector<ivec2> new_sites;
    for (int f = 0; f < nf; ++f)
    {
        if (rem[f])
            continue;
        g.appendFace(std::vector<int>(nfv.begin() + nfo[kept],
                                      nfv.begin() + nfo[kept + 1]));
        new_sites.push_back(m_face_sites[f]);
        ++kept;
    }
    g.finalize();
    m_geom = std::move(g);
    m_face_sites = std::move(new_sites);
    if (m_r_valid)
    {
        std::vector<float> r;
        for (int32_t i = 0; i < n_new; ++i)
            r.push_back(m_r_per_v[nvts[i]]);
        m_r_per_v = std::move(r);
    }
    if (!onlyHasInside())
    {
        m_vts_valid = new_valid;
        m_is_finite_v = new_finite;
    }
    else
    {
        m_vts_valid.assign(m_geom.numVts(), true);
        m_is_finite_v.assign(m_geom.numVts(), true);
    }
    recompute_flags();
}

void VoroInfo::computeEulerChar(eulerchar& ec, bool do_prune_first) const
{
    ensure_host();
    int V, E, F, C;
    if (do_prune_first)
    {
        std::vector<int> vi, ei, fi;
        getValidVts(vi);
        getValidEdges(ei);
        getValidFaces(fi);
        V = (int)vi.size();
        E = (int)ei.size();
        F = (int)fi.size();
        C = m_geom.compute_conn_cmpnts(&fi, 1);
    }
    else
    {
        V = (int)m_geom.numVts();
        E = (int)m_geom.numEdges();
        F = (int)m_geom.numFaces();
        C = m_geom.compute_conn_cmpnts();
    }
    ec = eulerchar(V, E, F, 0, C, V - E + F);
}

} // namespace voxelvoro
