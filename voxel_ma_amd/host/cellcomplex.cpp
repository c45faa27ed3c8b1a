#include "cellcomplex.h"

#include <stdexcept>

#include "vc_device.h"

cellcomplex::cellcomplex(const std::vector<point>& vts,
                         const std::vector<ivec2>& edges,
                         const std::vector<uTriFace>& faces,
                         bool perform_finalize)
{
    m_vts = vts;
    m_edges = edges;
    m_fv.reserve(3 * faces.size());
    m_face_off.reserve(faces.size() + 1);
    for (const auto& t : faces)
        appendFace(t);
    if (perform_finalize)
        finalize();
}

void cellcomplex::appendFace(const std::vector<int>& f)
{
    m_fv.insert(m_fv.end(), f.begin(), f.end());
    m_face_off.push_back((int64_t)m_fv.size());
    invalidate();
}

void cellcomplex::appendFace(const uTriFace& t)
{
    m_fv.push_back((int32_t)t[0]);
    m_fv.push_back((int32_t)t[1]);
    m_fv.push_back((int32_t)t[2]);
    m_face_off.push_back((int64_t)m_fv.size());
    invalidate();
}

std::vector<int32_t> cellcomplex::edgeArray() const
{
    std::vector<int32_t> a(2 * m_edges.size());
    for (size_t e = 0; e < m_edges.size(); ++e)
    {
        a[2 * e] = m_edges[e][0];
        a[2 * e + 1] = m_edges[e][1];
    }
    return a;
}

void cellcomplex::finalize()
{
    if (m_finalized)
        return;
    vc_ctx* ctx = vcdev::context();
    const int32_t nv = (int32_t)m_vts.size(), nf = (int32_t)numFaces();
    const std::vector<int32_t> ein = edgeArray();
    vc_cc* h = nullptr;
    vcdev::check(vc_cc_create(ctx, nv, (int32_t)m_edges.size(), ein.data(), nf,
                              m_face_off.data(), m_fv.data(), &h),
                 "vc_cc_create");
    adoptFinalized(std::shared_ptr<vc_cc>(h, [ctx](vc_cc* p) { vc_cc_release(ctx, p); }));
}

void cellcomplex::adoptFinalized(std::shared_ptr<vc_cc> dev)
{
    vc_ctx* ctx = vcdev::context();
    const int32_t nv = (int32_t)m_vts.size(), nf = (int32_t)numFaces();
    vc_cc* h = dev.get();
    int32_t dnv = 0, dnf = 0, ne = 0;
    int64_t ns = 0;
    vcdev::check(vc_cc_sizes(h, &dnv, &ne, &dnf, &ns), "vc_cc_sizes");
    if (dnv != nv || dnf != nf || ns != m_face_off.back())
        throw std::runtime_error("cellcomplex: the device complex has other sizes");
    m_dev = std::move(dev);
    std::vector<int32_t> eout(2 * (size_t)ne + 2);
    m_fe.assign((size_t)ns, 0);
    vcdev::check(vc_cc_get(ctx, h, eout.data(), m_fe.data()), "vc_cc_get");
    m_edges.resize(ne);
    for (int32_t e = 0; e < ne; ++e)
        m_edges[e] = ivec2(eout[2 * e], eout[2 * e + 1]);
    // host adjacency for the accessors: vertex -> edges in edge order,
    // edge -> faces in face order (the device holds the same lists)
    m_ve_off.assign((size_t)nv + 1, 0);
    m_ef_off.assign((size_t)ne + 1, 0);
    for (const auto& e : m_edges)
    {
        m_ve_off[e[0] + 1]++;
        m_ve_off[e[1] + 1]++;
    }
    for (int32_t id : m_fe)
        m_ef_off[id + 1]++;
    for (int32_t v = 0; v < nv; ++v)
        m_ve_off[v + 1] += m_ve_off[v];
    for (int32_t e = 0; e < ne; ++e)
        m_ef_off[e + 1] += m_ef_off[e];
    m_ve.resize(m_ve_off[nv]);
    m_ef.resize(m_ef_off[ne]);
    std::vector<int64_t> fv(m_ve_off.begin(), m_ve_off.end() - 1);
    for (int32_t e = 0; e < ne; ++e)
    {
        m_ve[fv[m_edges[e][0]]++] = e;
        m_ve[fv[m_edges[e][1]]++] = e;
    }
    std::vector<int64_t> ff(m_ef_off.begin(), m_ef_off.end() - 1);
    for (int32_t f = 0; f < nf; ++f)
        for (int64_t i = m_face_off[f]; i < m_face_off[f + 1]; ++i)
            m_ef[ff[m_fe[i]]++] = f;
    m_finalized = true;
}

void cellcomplex::getFaceVRep(int fi, std::vector<int>& f) const
{
    f.assign(m_fv.begin() + m_face_off[fi], m_fv.begin() + m_face_off[fi + 1]);
}

void cellcomplex::getFaceERep(int fi, std::vector<int>& f) const
{
    f.assign(m_fe.begin() + m_face_off[fi], m_fe.begin() + m_face_off[fi + 1]);
}

std::vector<int> cellcomplex::refCntPerVert() const
{
    std::vector<int> r(numVts());
    for (size_t v = 0; v < r.size(); ++v)
        r[v] = cntNbEdgesofVert((unsigned)v);
    return r;
}

std::vector<int> cellcomplex::refCntPerEdge() const
{
    std::vector<int> r(numEdges());
    for (size_t e = 0; e < r.size(); ++e)
        r[e] = cntNbFacesofEdge((unsigned)e);
    return r;
}

void cellcomplex::eulerChar(eulerchar& ec) const
{
    vc_ctx* ctx = vcdev::context();
    vc_cc_topo t{};
    if (m_finalized && m_dev)
        vcdev::check(vc_cc_topology(ctx, m_dev.get(), &t), "vc_cc_topology");
    else
    {
        // not finalized: the edges as given (the reference's const
        // overload, src/cellcomplex.cpp:493-502)
        const std::vector<int32_t> e = edgeArray();
        vcdev::check(vc_euler_char(ctx, (int32_t)m_vts.size(), e.data(),
                                   (int64_t)m_edges.size(), nullptr, 0, &t),
                     "vc_euler_char");
    }
    ec.V = (int)numVts();
    ec.E = (int)numEdges();
    ec.F = (int)numFaces();
    ec.T = 0;
    ec.euler = ec.V - ec.E + ec.F;
    ec.C = (int)t.C;
}

int cellcomplex::compute_conn_cmpnts() const
{
    if (!m_finalized || !m_dev)
    {
        cellcomplex c = *this;
        c.finalize();
        return c.compute_conn_cmpnts();
    }
    vc_cc_topo t{};
    vcdev::check(vc_cc_topology(vcdev::context(), m_dev.get(), &t), "vc_cc_topology");
    return (int)t.face_components;
}

int cellcomplex::compute_conn_cmpnts(const std::vector<int>* subset, int) const
{
    if (!m_finalized || !m_dev)
    {
        cellcomplex c = *this;
        c.finalize();
        return c.compute_conn_cmpnts(subset, 1);
    }
    if (!subset)
    {
        // every vertex, joined through the edges (src/cellcomplex.cpp:203-213)
        vc_cc_topo t{};
        vcdev::check(vc_cc_topology(vcdev::context(), m_dev.get(), &t),
                     "vc_cc_topology");
        return (int)t.C;
    }
    std::vector<uint8_t> in(numFaces(), 0);
    for (int f : *subset)
        if (f >= 0 && (size_t)f < in.size())
            in[f] = 1;
    int64_t n = 0;
    vcdev::check(vc_cc_subset_components(vcdev::context(), m_dev.get(), in.data(), &n),
                 "vc_cc_subset_components");
    return (int)n;
}
