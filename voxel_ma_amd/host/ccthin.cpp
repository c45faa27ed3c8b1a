#include "ccthin.h"

#include "vc_device.h"

void CellComplexThinning::release()
{
    if (m_dev)
        vc_ccthin_release(vcdev::context(), m_dev);
    m_dev = nullptr;
    m_cc_dev.reset();
    m_measures_sent = false;
}

void CellComplexThinning::reset()
{
    release();
    m_cc = nullptr;
    for (auto& m : m_measure)
        m.clear();
    for (auto& r : m_removed)
        r.clear();
    m_ref_edge.clear();
    m_host = false;
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
    m_measures_sent = false;
}

void CellComplexThinning::prune_host(float f_t, float l_t)
{
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
    m_host = true;
}

void CellComplexThinning::prune(float f_t, float l_t, bool)
{
    m_cc->finalize();
    vc_ctx* ctx = vcdev::context();
    const int32_t nv = (int32_t)m_cc->numVts(), ne = (int32_t)m_cc->numEdges(),
                  nf = (int32_t)m_cc->numFaces();
    if (m_cc->device() != m_cc_dev)
    {
        release();
        vcdev::check(vc_ccthin_create(ctx, m_cc->device().get(), &m_dev),
                     "vc_ccthin_create");
        m_cc_dev = m_cc->device();
    }
    if (!m_measures_sent)
    {
        std::vector<float> em = m_measure[EDGE], fm = m_measure[FACE];
        em.resize(ne, 0.0f);
        fm.resize(nf, 0.0f);
        vcdev::check(vc_ccthin_assign(ctx, m_dev, em.data(), fm.data()),
                     "vc_ccthin_assign");
        m_measures_sent = true;
    }
    const int rc = vc_ccthin_prune(ctx, m_dev, f_t, l_t, nullptr);
    if (rc == VC_ERR_INVALID)
    {
        // an edge with equal ends or a face listing a side twice
        prune_host(f_t, l_t);
        return;
    }
    vcdev::check(rc, "vc_ccthin_prune");
    std::vector<uint8_t> rv(nv), re(ne), rf(nf);
    m_ref_edge.assign(ne, 0);
    vcdev::check(vc_ccthin_removed(ctx, m_dev, rv.data(), re.data(), rf.data(),
                                   m_ref_edge.data()),
                 "vc_ccthin_removed");
    m_removed[VERTEX].assign(rv.begin(), rv.end());
    m_removed[EDGE].assign(re.begin(), re.end());
    m_removed[FACE].assign(rf.begin(), rf.end());
    m_host = false;
}

namespace
{
// which: 0 vertices, 1 isolated edges, 2 faces
void remained_ids(vc_ccthin* t, int which, std::vector<int>& ids)
{
    vc_ctx* ctx = vcdev::context();
    int64_t n[3] = {0, 0, 0};
    vcdev::check(vc_ccthin_remained(ctx, t, nullptr, 0, &n[0], nullptr, 0, &n[1],
                                    nullptr, 0, &n[2]),
                 "vc_ccthin_remained");
    std::vector<int32_t> out((size_t)n[which] + 1);
    int32_t* p[3] = {nullptr, nullptr, nullptr};
    p[which] = out.data();
    vcdev::check(vc_ccthin_remained(ctx, t, p[0], n[0], &n[0], p[1], n[1], &n[1],
                                    p[2], n[2], &n[2]),
                 "vc_ccthin_remained");
    ids.assign(out.begin(), out.begin() + n[which]);
}
} // namespace

void CellComplexThinning::getRemainedVts(std::vector<int>& ids) const
{
    ids.clear();
    if (!m_host)
    {
        remained_ids(m_dev, 0, ids);
        return;
    }
    for (size_t i = 0; i < m_cc->numVts(); ++i)
        if (!m_removed[VERTEX][i])
            ids.push_back((int)i);
}

void CellComplexThinning::getRemainedEdges(std::vector<int>& ids) const
{
    ids.clear();
    if (!m_host)
    {
        remained_ids(m_dev, 1, ids);
        return;
    }
    for (size_t i = 0; i < m_cc->numEdges(); ++i)
        if (!m_removed[EDGE][i] && m_ref_edge[i] == 0)
            ids.push_back((int)i);
}

void CellComplexThinning::getRemainedFaces(std::vector<int>& ids) const
{
    ids.clear();
    if (!m_host)
    {
        remained_ids(m_dev, 2, ids);
        return;
    }
    for (size_t i = 0; i < m_cc->numFaces(); ++i)
        if (!m_removed[FACE][i])
            ids.push_back((int)i);
}

void CellComplexThinning::remainingCC(std::vector<point>& vts,
                                      std::vector<ivec2>& edges,
                                      std::vector<uTriFace>& tris,
                                      std::vector<int>* from_fi) const
{
    if (!m_host)
    {
        // on the device, sized by upper bounds so one call fills everything
        vc_ctx* ctx = vcdev::context();
        const auto& fo = m_cc->faceOffsets();
        int64_t tcap = 0;
        for (size_t f = 0; f + 1 < fo.size(); ++f)
            tcap += fo[f + 1] - fo[f] >= 3 ? fo[f + 1] - fo[f] - 2 : 0;
        const int64_t vcap = (int64_t)m_cc->numVts(), ecap = (int64_t)m_cc->numEdges();
        std::vector<int32_t> vi((size_t)vcap + 1), e(2 * (size_t)ecap + 2),
            t(3 * (size_t)tcap + 3), ff((size_t)tcap + 1);
        int64_t nv = 0, ne = 0, nt = 0;
        vcdev::check(vc_ccthin_remaining(ctx, m_dev, vi.data(), vcap, &nv, e.data(),
                                         ecap, &ne, t.data(),
                                         from_fi ? ff.data() : nullptr, tcap, &nt),
                     "vc_ccthin_remaining");
        vts.resize((size_t)nv);
        for (int64_t i = 0; i < nv; ++i)
            vts[i] = m_cc->getVert(vi[i]);
        edges.resize((size_t)ne);
        for (int64_t i = 0; i < ne; ++i)
            edges[i] = ivec2(e[2 * i], e[2 * i + 1]);
        tris.resize((size_t)nt);
        for (int64_t i = 0; i < nt; ++i)
            tris[i] = uTriFace((unsigned)t[3 * i], (unsigned)t[3 * i + 1],
                               (unsigned)t[3 * i + 2]);
        if (from_fi)
            from_fi->assign(ff.begin(), ff.begin() + nt);
        return;
    }
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
