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
    for (const auto& t : faces)
        appendFace(t);
    if (perform_finalize)
        finalize();
}

void cellcomplex::appendFace(const std::vector<int>& f)
{
    m_fv.insert(m_fv.end(), f.begin(), f.end());
    m_face_off.push_back((int64_t)m_fv.size());
    m_finalized = false;
}

void cellcomplex::appendFace(const uTriFace& t)
{
    appendFace(std::vector<int>{(int)t[0], (int)t[1], (int)t[2]});
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
    const int32_t nv = (int32_t)m_vts.size(), nf = (int32_t)numFaces();
    std::vector<int32_t> ein = edgeArray();
    std::vector<int32_t> eout(ein.size() + 2 * m_fv.size());
    std::vector<int32_t> fe(m_fv.size());
    int32_t ne = 0;
    vcdev::check(vc_cc_finalize(nv, (int32_t)m_edges.size(), ein.data(), nf,
                                m_face_off.data(), m_fv.data(), eout.data(),
                                (int32_t)(eout.size() / 2), &ne, fe.data()),
                 "vc_cc_finalize");
    m_edges.resize(ne);
    for (int32_t e = 0; e < ne; ++e)
        m_edges[e] = ivec2(eout[2 * e], eout[2 * e + 1]);
    m_fe.swap(fe);
    // adjacency: vertex -> edges in edge order, edge -> faces in face order
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
    const int C = util::findNumConnComponents(m_vts, m_edges);
    ec.V = (int)numVts();
    ec.E = (int)numEdges();
    ec.F = (int)numFaces();
    ec.T = 0;
    ec.euler = ec.V - ec.E + ec.F;
    ec.C = C;
}

int cellcomplex::compute_conn_cmpnts() const
{
    const size_t nf = numFaces();
    std::vector<bool> fvis(nf, false), evis(numEdges(), false);
    std::vector<int> stack, fe;
    int ncc = 0;
    for (size_t s = 0; s < nf; ++s)
    {
        if (fvis[s])
            continue;
        ++ncc;
        stack.assign(1, (int)s);
        while (!stack.empty())
        {
            const int f = stack.back();
            stack.pop_back();
            if (fvis[f])
                continue;
            fvis[f] = true;
            getFaceERep(f, fe);
            for (int e : fe)
            {
                if (evis[e])
                    continue;
                evis[e] = true;
                for (int k = 0; k < cntNbFacesofEdge(e); ++k)
                    if (!fvis[nbFaceofEdge(e, k)])
                        stack.push_back(nbFaceofEdge(e, k));
            }
        }
    }
    return ncc;
}
