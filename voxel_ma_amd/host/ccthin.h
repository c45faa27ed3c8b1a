// ccthin.h -- CellComplexThinning, the reference's measure-thresholded
// simple-pair collapse of a 2-complex (include/ccthin.h:13-167).  prune()
// runs on the device over the complex cellcomplex::finalize left there
// (vc_ccthin_prune, include/vc_cc.h), in the reference's FIFO order, and so
// do remainingCC and getRemained*; a complex the device collapse refuses (an
// edge with equal ends, a face listing a side twice) is collapsed on the
// host in the same order (vc_cc_prune).
#pragma once

#include <memory>
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
    ~CellComplexThinning() { release(); }
    CellComplexThinning(const CellComplexThinning&) = delete;
    CellComplexThinning& operator=(const CellComplexThinning&) = delete;
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
    void release();
    void prune_host(float f_t, float l_t);

    cellcomplex* m_cc = nullptr;
    std::vector<float> m_measure[3];
    std::vector<bool> m_removed[3];
    std::vector<int> m_ref_edge;
    // the collapse state on the device, over the complex it was made for
    std::shared_ptr<vc_cc> m_cc_dev;
    vc_ccthin* m_dev = nullptr;
    bool m_measures_sent = false;
    bool m_host = false; // the last prune ran on the host
    float m_face_degen_thresh = 0.0f;
    int m_cmpnt_num_faces_thresh = 0;
};
