// geomalgo.h -- the helpers of the reference's util:: namespace that the
// thinning path uses (include/geomalgo.h:79-141, src/geomalgo.cpp:13-235).
#pragma once

#include <vector>

#include "commondefs.h"

namespace util
{
using std::vector;

// sorted vertex pair (include/geomalgo.h:79-82)
inline ivec2 makeEdge(int i, int j) { return i > j ? ivec2(j, i) : ivec2(i, j); }

// polygon -> its sides as sorted pairs (include/geomalgo.h:129-138)
inline void traceFace(const vector<int>& f_of_vts, vector<ivec2>& f_of_edges)
{
    f_of_edges.resize(f_of_vts.size());
    for (size_t i = 0; i < f_of_vts.size(); ++i)
        f_of_edges[i] =
            makeEdge(f_of_vts[i], f_of_vts[(i + 1) % f_of_vts.size()]);
}

// zig-zag fan: (last, 0, 1), (last, 1, last-1), ... (src/geomalgo.cpp:13-35)
void simpleTriangulate(const vector<int>& poly, vector<uTriFace>& tris);
// renumber vertex ids by their position in v_l (src/geomalgo.cpp:215-235)
void compactify(const vector<int>& v_l, vector<ivec2>& e_l,
                vector<uTriFace>& f_l);
// connected components over a vertex adjacency table; vertices already
// marked visited are skipped (src/geomalgo.cpp:67-119)
int findNumConnComponents(const vector<vector<int>>& v_v_adj, int n_vts,
                          vector<bool>& visited);
int findNumConnComponents(const vector<point>& vts, const vector<ivec2>& edges);
// V, E, F, C, chi of a triangle soup / of vertices + edges + triangles
// (src/geomalgo.cpp:120-214)
void computeEulerChar(const vector<uTriFace>& tris, eulerchar& ec);
void computeEulerChar(const vector<point>& vts, const vector<ivec2>& edges,
                      const vector<uTriFace>& tris, eulerchar& ec);
} // namespace util
