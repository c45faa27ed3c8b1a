// surfacing.h -- Surfacer, the reference's boundary-site and cube-complex
// Euler extractor (include/surfacing.h:17-69), computed on the GPU
// (vc_boundary_extract) with the reference's exact output order.
#pragma once

#include <memory>
#include <vector>

#include "Volume3DScalar.h"
#include "commondefs.h"

enum class SurfacerErrCode
{
    FAILURE,
    SUCCESS,
    EMPTY_VOL,
    INVALID_VOL_TYPE
};

class Surfacer
{
public:
    Surfacer() = default;
    explicit Surfacer(std::shared_ptr<Volume3DScalar> vol) : m_vol(std::move(vol)) {}

    // quads: 4 indices into vts tracing each boundary face
    SurfacerErrCode extractBoundaryVtsAndFaces(const std::shared_ptr<Volume3DScalar>& vol,
                                               std::vector<point>& vts,
                                               std::vector<ivec4>& faces,
                                               bool need_euler);
    // triangles: each quad split as (0,1,2), (2,3,0) (src/surfacing.cpp:591-595)
    SurfacerErrCode extractBoundaryVtsAndFaces(const std::shared_ptr<Volume3DScalar>& vol,
                                               std::vector<point>& vts,
                                               std::vector<uTriFace>& faces,
                                               eulerchar* euler = nullptr);
    SurfacerErrCode extractBoundaryVts(const std::shared_ptr<Volume3DScalar>& vol,
                                       std::vector<point>& vts);
    // Euler characteristic of the volume's cube complex; vts / faces must be
    // this Surfacer's own output for vol (they are not re-read)
    int computeEulerChar(const std::vector<point>& vts,
                         const std::vector<ivec4>& faces,
                         const std::shared_ptr<Volume3DScalar>& vol,
                         eulerchar& euler) const;

private:
    SurfacerErrCode run(const std::shared_ptr<Volume3DScalar>& vol,
                        std::vector<point>* vts, std::vector<ivec4>* quads,
                        eulerchar* euler) const;
    std::shared_ptr<Volume3DScalar> m_vol;
};
