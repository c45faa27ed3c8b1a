// voroinfo.h -- VoroInfo with the reference's names for the Voronoi-side
// steps of the pipeline (include/voroinfo.h, src/voroinfo.cpp): inside
// tagging, lambda measures, inside extraction with measures, close-vertex
// merging.  The diagram is given as data (setGeometry: vertices, face
// polygons, the two sites of every face, site positions, finite flags), not
// built from TetGen output, or kept resident on the device (setDevice, as
// computeVD does: include/vc_vdiag.h), where the steps of vol2ma run without
// copying it; the host geometry is then filled in from the device on the
// first call that needs it.  Tagging, measures, extraction and the
// nothing-merges case of merging run on the device (include/vc_voro.h).
#pragma once

#include <memory>
#include <vector>

#include "Volume3DScalar.h"
#include "cellcomplex.h"
#include "commondefs.h"

struct vc_vdiag;
struct vc_cc;

namespace MeasureForMA
{
enum meassuretype
{
    LAMBDA = 0
};
// |a - b| as trimesh::dist (include/measureforMA_imp.h:1-4)
float lambdaForFace(const point& a, const point& b);
} // namespace MeasureForMA

namespace voxelvoro
{

class VoroInfo
{
public:
    VoroInfo() = default;
    // vts: Voronoi vertices; faces: vertex polygons; face_sites: the two
    // sites of each face; sites: site positions; finite_v: empty = all
    // finite.  The complex is finalized (edges derived as cellcomplex).
    void setGeometry(const std::vector<point>& vts,
                     const std::vector<std::vector<int>>& faces,
                     const std::vector<ivec2>& face_sites,
                     const std::vector<point>& sites,
                     const std::vector<bool>& finite_v = {});
    const cellcomplex& geom() const
    {
        ensure_host();
        return m_geom;
    }
    const std::vector<ivec2>& faceSites() const;
    const std::vector<point>& getSites() const { return getSitesPosition(); }

    // the diagram resident on the device (computeVD); copies of a VoroInfo
    // share it until one of them changes it
    void setDevice(std::shared_ptr<vc_vdiag> d);
    bool onDevice() const { return (bool)m_dev; }
    size_t numVts() const;
    size_t numEdges() const;
    size_t numFaces() const;
    // position of vertex vi (the resident diagram's, fetched once)
    const point& vertex(int vi) const;
    // the finalized complex of the output of the last
    // extractInsideWithMeasure on the device, handed over once (else null)
    std::shared_ptr<vc_cc> takeInsideComplex() const;

    // src/voroinfo.cpp:447-496
    bool tagVert(const point& p, const std::shared_ptr<Volume3DScalar>& vol) const;
    bool tagVtsUsingUniformVol(const std::shared_ptr<Volume3DScalar>& vol);
    float getInsidePartSize() const;
    void set_only_inside(bool yesno) { m_only_inside = yesno ? 1 : 0; }
    bool onlyHasInside() const { return m_only_inside > 0; }

    bool isVertexValid(int vi) const;
    bool isEdgeValid(int ei) const;
    bool isFaceValid(int fi) const;
    void getValidVts(std::vector<int>& idx) const;
    void getValidEdges(std::vector<int>& idx) const;
    void getValidFaces(std::vector<int>& idx) const;
    // include/voroinfo_imp.h:57-60: a face is finite iff all its vertices
    // are (the diagram loader's per-face flag)
    bool isFaceFinite(int fi) const;

    // src/voroinfo.cpp:1432-1574 (LAMBDA)
    void computeVertexMeasure(MeasureForMA::meassuretype tp,
                              const std::vector<int>& vts_indices,
                              std::vector<float>& vts_msure) const;
    void computeEdgesMeasure(MeasureForMA::meassuretype tp,
                             const std::vector<int>& edges_indices,
                             std::vector<float>& edges_msure) const;
    void computeFacesMeasure(MeasureForMA::meassuretype tp,
                             const std::vector<int>& faces_indices,
                             std::vector<float>& faces_msure) const;

    // src/voroinfo.cpp:624-729
    void extractInsideWithMeasure(MeasureForMA::meassuretype tp,
                                  std::vector<point>& out_vts,
                                  std::vector<ivec2>& out_edges,
                                  std::vector<uTriFace>& out_tris,
                                  std::vector<int>& from_fi,
                                  std::vector<float>& vts_msure,
                                  std::vector<float>& edges_msure,
                                  std::vector<float>& faces_msure) const;

    // src/voroinfo.cpp:1064-1384; the surviving edges are appended in
    // ascending order (the reference iterates an unordered_set); radii, when
    // set, follow the seeds (:1176-1188)
    void mergeCloseVts(float eps, bool only_inside);

    // radius per vertex: distance to the first site of the last face holding
    // it (src/voroinfo.cpp:290-325), computed on the device
    void computeRadii();
    bool radiiValid() const { return m_r_valid; }
    const std::vector<float>& getRadii() const;
    ivec2 getSitesOfFace(int fi) const { return faceSites()[fi]; }
    const std::vector<point>& getSitesPosition() const;

    // src/voroinfo.cpp:1386-1430: V, E, F of the valid part (do_prune_first)
    // or of the whole diagram; C = components of the valid faces' vertices
    // through their sides, or the face components of the whole complex
    // (cellcomplex::compute_conn_cmpnts, on the device)
    void computeEulerChar(eulerchar& ec, bool do_prune_first) const;

    // valid vertices, the valid triangles of every face's triangulation and
    // their sides, sorted (src/voroinfo.cpp:573-620)
    void getPartInside(std::vector<int>& vts_indices, std::vector<ivec2>& edges,
                       std::vector<uTriFace>& tri_faces) const;

private:
    void all_measures(std::vector<float>& lv, std::vector<float>& le,
                      std::vector<float>& lf) const;
    void recompute_flags();
    // the host geometry from the resident diagram (which is then dropped)
    void materialize();
    void ensure_host() const
    {
        if (m_dev)
            const_cast<VoroInfo*>(this)->materialize();
    }
    // before changing a resident diagram another VoroInfo shares
    void own_device();
    void drop_device_copies() const;

    cellcomplex m_geom;
    std::vector<ivec2> m_face_sites;
    std::vector<point> m_site_positions;
    std::vector<bool> m_vts_valid, m_is_finite_v, m_edge_valid, m_face_valid;
    bool m_tagged = false;
    int m_only_inside = -1;
    point m_bb_min, m_bb_max;
    bool m_bb_valid = false;
    std::vector<float> m_r_per_v;
    bool m_r_valid = false;
    // resident diagram and the host copies fetched from it
    std::shared_ptr<vc_vdiag> m_dev;
    mutable std::vector<point> m_dev_vts;
    mutable bool m_got_vts = false, m_got_fs = false, m_got_sites = false,
                 m_got_radii = false;
    mutable std::shared_ptr<vc_cc> m_inside_cc;
};

} // namespace voxelvoro
