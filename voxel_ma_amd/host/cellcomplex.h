// cellcomplex.h -- the reference's 2-complex (include/cellcomplex.h:23-237):
// vertices, edges and polygon faces, finalize() to extract the face sides
// and build vertex->edge / edge->face adjacency.  finalize runs on the
// device (vc_cc_create, include/vc_cc.h) with the reference's edge
// numbering and keeps the finalized complex resident there, where the
// collapse (CellComplexThinning) and the topology checks use it.
#pragma once

#include <memory>
#include <vector>

#include "../../include/vc_cc.h"
#include "commondefs.h"
#include "geomalgo.h"

class cellcomplex
{
public:
    cellcomplex() = default;
    cellcomplex(const std::vector<point>& vts, const std::vector<ivec2>& edges,
                const std::vector<uTriFace>& faces, bool perform_finalize = true);

    void appendVert(const point& v)
    {
        m_vts.push_back(v);
        invalidate();
    }
    void appendEdge(const ivec2& e)
    {
        m_edges.push_back(e);
        invalidate();
    }
    void appendFace(const std::vector<int>& f_of_vts);
    void appendFace(const uTriFace& f);
    void finalize();
    bool isFinalized() const { return m_finalized; }

    size_t numVts() const { return m_vts.size(); }
    size_t numEdges() const { return m_edges.size(); }
    size_t numFaces() const { return m_face_off.size() - 1; }
    const point& getVert(size_t i) const { return m_vts[i]; }
    const ivec2& getEdge(size_t i) const { return m_edges[i]; }
    void getFaceVRep(int fi, std::vector<int>& f) const;
    void getFaceERep(int fi, std::vector<int>& f) const;
    int cntNbEdgesofVert(unsigned vi) const { return (int)(m_ve_off[vi + 1] - m_ve_off[vi]); }
    int cntNbFacesofEdge(unsigned ei) const { return (int)(m_ef_off[ei + 1] - m_ef_off[ei]); }
    int nbEdgeofVert(unsigned vi, int k) const { return m_ve[m_ve_off[vi] + k]; }
    int nbFaceofEdge(unsigned ei, int k) const { return m_ef[m_ef_off[ei] + k]; }
    std::vector<int> refCntPerVert() const;
    std::vector<int> refCntPerEdge() const;
    // chi = V - E + F with C = connected components over the edges
    // (src/cellcomplex.cpp:493-512), on the device
    void eulerChar(eulerchar& ec) const;
    // connected components of faces through shared edges
    // (src/cellcomplex.cpp:74-165), on the device
    int compute_conn_cmpnts() const;
    // with a face subset and the tag (src/cellcomplex.cpp:166-219): the
    // vertices of the subset's faces joined by their polygon sides
    int compute_conn_cmpnts(const std::vector<int>* subset_faces, int tag) const;

    // the finalized complex on the device (finalize() first); shared by
    // copies of this complex
    const std::shared_ptr<vc_cc>& device() const { return m_dev; }
    // finalize with a complex already finalized on the device from this
    // complex's vertices, edges and faces (the inside part extracted on the
    // device), instead of creating it again
    void adoptFinalized(std::shared_ptr<vc_cc> dev);

    // flat views for the C ABI
    const std::vector<int64_t>& faceOffsets() const { return m_face_off; }
    const std::vector<int32_t>& faceEdgeIds() const { return m_fe; }
    std::vector<int32_t> edgeArray() const;

private:
    void invalidate()
    {
        m_finalized = false;
        m_dev.reset();
    }

    std::vector<point> m_vts;
    std::vector<ivec2> m_edges;
    std::vector<int64_t> m_face_off{0};
    std::vector<int32_t> m_fv, m_fe;
    std::vector<int64_t> m_ve_off, m_ef_off;
    std::vector<int32_t> m_ve, m_ef;
    bool m_finalized = false;
    std::shared_ptr<vc_cc> m_dev;
};
