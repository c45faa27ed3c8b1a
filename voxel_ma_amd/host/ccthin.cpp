#include "ccthin.h"

#include "vc_device.h"

void CellComplexThinning::reset()
{
    m_cc = nullptr;
    for (auto& m : m_measure)
        m.clear();
    for (auto& r : m_removed)
        r.clear();
    m_ref_edge.clear();
}

void CellComplexThinning::preprocess()
{
    m_removed[FACE].assign(m_cc->numFaces(), false);
    m_removed[EDGE].assign(m_cc->numEdges(), false);
    m_removed[VERTEX].assign(m_cc->numVts(), false);
}

void CellComplexThinning::assignElementValues(const std::vector<float>& v,
                                              const std::vector<float>& e,
                                              const std::vector<float>& f)
{
    m_measure[VERTEX] = v;
    m_measure[EDGE] = e;
    m_measure[FACE] = f;
}

void CellComplexThinning::prune(float f_t, float l_t, bool)
{
    m_cc->finalize();
    const int32_t nv = (int32_t)m_cc->numVts(), ne = (int32_t)m_cc->numEdges(),
                  nf = (int32_t)m_cc->numFaces();
    const std::vector<int32_t> edges = m_cc->edgeArray();
    vc_cc_view view{nv, ne, nf, edges.data(), m_cc->faceOffsets().data(),
                    m_cc->faceEdgeIds().data()};
    std::vector<uint8_t> rv(nv), re(ne), rf(nf);
    m_ref_edge.assign(ne, 0);
    std::vector<float> em = m_measure[EDGE], fm = m_measure[FACE];
    em.resize(ne, 0.0f);
    fm.resize(nf, 0.0f);
    vcdev::check(vc_cc_prune(&view, em.data(), fm.data(), f_t, l_t, rv.data(),
                             re.data(), rf.data(), m_ref_edge.data()),
                 "vc_cc_prune");
    m_removed[VERTEX].assign(rv.begin(), rv.end());
    m_removed[EDGE].assign(re.begin(), re.end());
    m_removed[FACE].assign(rf.begin(), rf.end());
}

void CellComplexThinning::getRemainedVts(std::vector<int>& ids) const
{
    ids.clear();
    for (size_t i = 0; i < m_cc->numVts(); ++i)
        if (!m_removed[VERTEX][i])
            ids.push_back((int)i);
}

void CellComplexThinning::getRemainedEdges(std::vector<int>& ids) const
{
    ids.clear();
    for (size_t i = 0; i < m_cc->numEdges(); ++i)
        if (!m_removed[EDGE][i] && m_ref_edge[i] == 0)
            ids.push_back((int)i);
}

void CellComplexThinning::getRemainedFaces(std::vector<int>& ids) const
{
    ids.clear();
    for (size_t i = 0; i < m_cc->numFaces(); ++i)
        if (!m_removed[FACE][i])
            ids.push_back((int)i);
}

void CellComplexThinning::remainingCC(std::vector<point>& vts,
                                      std::vector<ivec2>& edges,
                                      std::vector<uTriFace>& tris,
                                      std::vector<int>* from_fi) const
{
    std::vector<int> ids;
    getRemainedEdges(ids);
    edges.clear();
    for (int e : ids)
        edges.push_back(m_cc->getEdge(e));
    getRemainedFaces(ids);
    tris.clear();
    if (from_fi)
        from_fi->clear();
    std::vector<int> f;
    std::vector<uTriFace> ft;
    for (int fi : ids)
    {
        m_cc->getFaceVRep(fi, f);
        util::simpleTriangulate(f, ft);
        tris.insert(tris.end(), ft.begin(), ft.end());
        if (from_fi)
            from_fi->insert(from_fi->end(), ft.size(), fi);
    }
    getRemainedVts(ids);
    vts.clear();
    for (int v : ids)
        vts.push_back(m_cc->getVert(v));
    util::compactify(ids, edges, tris);
}

cellcomplex CellComplexThinning::remainingCC()
{
    std::vector<point> v;
    std::vector<ivec2> e;
    std::vector<uTriFace> t;
    remainingCC(v, e, t, nullptr);
    return cellcomplex(v, e, t);
}
