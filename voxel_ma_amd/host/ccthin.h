// ccthin.h -- CellComplexThinning, the reference's measure-thresholded
// simple-pair collapse of a 2-complex (include/ccthin.h:13-167).  prune()
// runs through the C ABI (vc_cc_prune) in the reference's FIFO order.
#pragma once

#include <vector>

#include "cellcomplex.h"

class CellComplexThinning
{
public:
    enum celltype
    {
        FACE = 1,
        EDGE = 0,
        VERTEX = 2
    };
    CellComplexThinning() { reset(); }
    // non-owning, as in the reference (the caller keeps the complex alive)
    void setup(cellcomplex* cc) { m_cc = cc; }
    void reset();
    void preprocess();
    const std::vector<bool>& getEdgeRemovedTags() const { return m_removed[EDGE]; }
    const std::vector<bool>& getFaceRemovedTags() const { return m_removed[FACE]; }
    // not removed vertices / isolated not removed edges / not removed faces
    void getRemainedVts(std::vector<int>& ids) const;
    void getRemainedEdges(std::vector<int>& ids) const;
    void getRemainedFaces(std::vector<int>& ids) const;
    // measures are copied (src/ccthin.cpp:178-189)
    void assignElementValues(const std::vector<float>& v,
                             const std::vector<float>& e,
                             const std::vector<float>& f);
    void setPolyFaceDegenerateThresh(float t) { m_face_degen_thresh = t; }
    void setComponentFaceNumberThresh(int t) { m_cmpnt_num_faces_thresh = t; }
    // face-edge pairs with face measure < f_t, edge-vertex pairs with edge
    // measure < l_t; every call starts from the full complex
    void prune(float f_t, float l_t, bool remove_small_components);
    // remaining complex, faces triangulated (util::simpleTriangulate), vertex
    // ids compacted (src/ccthin.cpp:426-512)
    cellcomplex remainingCC();
    void remainingCC(std::vector<point>& vts, std::vector<ivec2>& edges,
                     std::vector<uTriFace>& tris,
                     std::vector<int>* from_fi = nullptr) const;

private:
    cellcomplex* m_cc = nullptr;
    std::vector<float> m_measure[3];
    std::vector<bool> m_removed[3];
    std::vector<int> m_ref_edge;
    float m_face_degen_thresh = 0.0f;
    int m_cmpnt_num_faces_thresh = 0;
};
