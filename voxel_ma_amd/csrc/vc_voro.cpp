// vc_voro.cpp -- inside extraction and close-vertex merging of a Voronoi
// 2-complex (host side of include/vc_voro.h).  Both are order-defined walks
// over the complex (the output numbering is the reference's discovery
// order), so they run on the host, sequentially.
#include <algorithm>
#include <cstring>
#include <queue>
#include <unordered_set>
#include <vector>

#include "vc_internal.h"
#include "../../include/vc_voro.h"

// float distances as the reference's (separately rounded products and sums)
#pragma clang fp contract(off)

namespace
{

struct PairHash
{
    size_t operator()(uint64_t k) const { return (size_t)(k * 0x9E3779B97F4A7C15ull); }
};

inline uint64_t edge_key(int32_t a, int32_t b)
{
    if (a > b)
        std::swap(a, b);
    return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}

// zig-zag triangulation of a polygon: (p[i], p[j], p[j+1]) and
// (p[i], p[j], p[i-1]) alternately, i from the last vertex down, j from the
// first up (src/geomalgo.cpp:13-35)
void zigzag(const int32_t* p, int n, std::vector<int32_t>& tris)
{
    tris.clear();
    int i = n - 1, j = 0;
    bool inc_j = true;
    while (i - 1 > j)
    {
        const int k = inc_j ? j + 1 : i - 1;
        tris.push_back(p[i]);
        tris.push_back(p[j]);
        tris.push_back(p[k]);
        if (inc_j)
            ++j;
        else
            --i;
        inc_j = !inc_j;
    }
}

bool face_valid(const int64_t* face_off, const int32_t* face_vts,
                const uint8_t* v_valid, int32_t f)
{
    for (int64_t j = face_off[f]; j < face_off[f + 1]; ++j)
        if (!v_valid[face_vts[j]])
            return false;
    return true;
}

int check_complex(int32_t nv, int32_t ne, int32_t nf, const int32_t* edges,
                  const int64_t* face_off, const int32_t* face_vts)
{
    if (nv < 0 || ne < 0 || nf < 0 || (ne > 0 && !edges) ||
        (nf > 0 && (!face_off || !face_vts)))
        return VC_ERR_INVALID;
    for (int64_t i = 0; i < 2LL * ne; ++i)
        if (edges[i] < 0 || edges[i] >= nv)
            return VC_ERR_INVALID;
    if (nf > 0 && face_off[0] != 0)
        return VC_ERR_INVALID;
    for (int32_t f = 0; f < nf; ++f)
    {
        if (face_off[f + 1] < face_off[f])
            return VC_ERR_INVALID;
        for (int64_t j = face_off[f]; j < face_off[f + 1]; ++j)
            if (face_vts[j] < 0 || face_vts[j] >= nv)
                return VC_ERR_INVALID;
    }
    return VC_OK;
}

} // namespace

extern "C" {

int vc_voro_extract_inside(int32_t nv, int32_t ne, int32_t nf,
                           const int32_t* edges, const int64_t* face_off,
                           const int32_t* face_vts, const uint8_t* v_valid,
                           const float* lam_e, const float* lam_f,
                           int32_t* vts_idx, int64_t vts_cap, int64_t* n_vts,
                           int32_t* edges_out, float* em, int64_t edge_cap,
                           int64_t* n_edges, int32_t* tris, int32_t* from_fi,
                           float* fm, int64_t tri_cap, int64_t* n_tris)
{
    int rc = check_complex(nv, ne, nf, edges, face_off, face_vts);
    if (rc != VC_OK || (nv > 0 && !v_valid) || (ne > 0 && !lam_e) ||
        (nf > 0 && !lam_f) || !n_vts || !n_edges || !n_tris)
        return VC_ERR_INVALID;
    // valid vertices (ascending) and their compact ids
    std::vector<int32_t> vid, new_id((size_t)nv, 0);
    for (int32_t v = 0; v < nv; ++v)
        if (v_valid[v])
        {
            new_id[(size_t)v] = (int32_t)vid.size();
            vid.push_back(v);
        }
    // valid edges as stored, then triangulation sides not seen before
    std::vector<int32_t> e_out;
    std::vector<float> e_ms;
    std::unordered_set<uint64_t, PairHash> seen;
    for (int32_t e = 0; e < ne; ++e)
    {
        const int32_t a = edges[2 * e], b = edges[2 * e + 1];
        if (!(v_valid[a] && v_valid[b]))
            continue;
        e_out.push_back(a);
        e_out.push_back(b);
        e_ms.push_back(lam_e[e]);
        // exact-pair membership, as stored
        seen.insert(((uint64_t)(uint32_t)a << 32) | (uint32_t)b);
    }
    std::vector<int32_t> t_out, t_from, ft;
    std::vector<float> t_ms;
    for (int32_t f = 0; f < nf; ++f)
    {
        if (!face_valid(face_off, face_vts, v_valid, f))
            continue;
        zigzag(face_vts + face_off[f], (int)(face_off[f + 1] - face_off[f]), ft);
        for (size_t t = 0; t < ft.size(); t += 3)
        {
            t_ms.push_back(lam_f[f]);
            t_out.insert(t_out.end(), ft.begin() + t, ft.begin() + t + 3);
            t_from.push_back(f);
            for (int s = 0; s < 3; ++s)
            {
                int32_t a = ft[t + s], b = ft[t + (s + 1) % 3];
                if (a > b)
                    std::swap(a, b);
                const uint64_t k = ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
                if (seen.insert(k).second)
                {
                    e_out.push_back(a);
                    e_out.push_back(b);
                    e_ms.push_back(lam_f[f]);
                }
            }
        }
    }
    *n_vts = (int64_t)vid.size();
    *n_edges = (int64_t)e_ms.size();
    *n_tris = (int64_t)t_ms.size();
    if ((vts_idx && vts_cap < *n_vts) || (edges_out && edge_cap < *n_edges) ||
        (tris && tri_cap < *n_tris))
        return VC_ERR_CAPACITY;
    // renumber into the valid-vertex order (util::compactify)
    if (vts_idx)
        std::copy(vid.begin(), vid.end(), vts_idx);
    if (edges_out)
        for (size_t i = 0; i < e_out.size(); ++i)
            edges_out[i] = new_id[(size_t)e_out[i]];
    if (em)
        std::copy(e_ms.begin(), e_ms.end(), em);
    if (tris)
        for (size_t i = 0; i < t_out.size(); ++i)
            tris[i] = new_id[(size_t)t_out[i]];
    if (from_fi)
        std::copy(t_from.begin(), t_from.end(), from_fi);
    if (fm)
        std::copy(t_ms.begin(), t_ms.end(), fm);
    return VC_OK;
}

int vc_voro_merge_close(int32_t nv, int32_t ne, int32_t nf, const float* vts,
                        const int32_t* edges, const int64_t* face_off,
                        const int32_t* face_vts, const int64_t* ve_off,
                        const int32_t* ve, const uint8_t* v_valid, float eps,
                        int only_inside, int32_t* old_to_new, int32_t* new_vts,
                        int32_t* n_new_vts, int64_t* new_face_off,
                        int32_t* new_face_vts, uint8_t* remove_face,
                        int32_t* new_edges, int64_t edge_cap,
                        int64_t* n_new_edges, uint8_t* keep_edge)
{
    int rc = check_complex(nv, ne, nf, edges, face_off, face_vts);
    if (rc != VC_OK || (nv > 0 && (!vts || !ve_off || !old_to_new || !new_vts)) ||
        (only_inside && nv > 0 && !v_valid) || !n_new_vts || !n_new_edges ||
        (nf > 0 && (!new_face_off || !new_face_vts || !remove_face)) ||
        (ne > 0 && !keep_edge) || !(eps >= 0.0f))
        return VC_ERR_INVALID;
    for (int64_t i = 0; nv > 0 && i < ve_off[nv]; ++i)
        if (ve[i] < 0 || ve[i] >= ne)
            return VC_ERR_INVALID;
    const float sq_r = eps * eps;
    auto dist2 = [&](int32_t a, int32_t b) {
        // trimesh::dist2 order, float
        const float d0 = vts[3 * b] - vts[3 * a];
        const float d1 = vts[3 * b + 1] - vts[3 * a + 1];
        const float d2 = vts[3 * b + 2] - vts[3 * a + 2];
        float q = d0 * d0;
        q += d1 * d1;
        q += d2 * d2;
        return q;
    };
    std::vector<uint8_t> visited((size_t)nv, 0);
    for (int32_t v = 0; v < nv; ++v)
        old_to_new[v] = -1;
    int32_t nn = 0;
    std::queue<int32_t> q;
    for (int32_t vi = 0; vi < nv; ++vi)
    {
        if (visited[vi])
            continue;
        if (only_inside && !v_valid[vi])
            continue;
        new_vts[nn] = vi;
        q.push(vi);
        visited[vi] = 1;
        old_to_new[vi] = nn;
        while (!q.empty())
        {
            const int32_t cur = q.front();
            q.pop();
            for (int64_t i = ve_off[cur]; i < ve_off[cur + 1]; ++i)
            {
                const int32_t e = ve[i];
                const int32_t ni =
                    edges[2 * e] == cur ? edges[2 * e + 1] : edges[2 * e];
                if (!visited[ni] && dist2(vi, ni) <= sq_r)
                {
                    q.push(ni);
                    visited[ni] = 1;
                    old_to_new[ni] = nn;
                }
            }
        }
        ++nn;
    }
    *n_new_vts = nn;
    std::vector<uint64_t> eset;
    for (int32_t e = 0; e < ne; ++e)
    {
        keep_edge[e] = 0;
        const int32_t a = edges[2 * e], b = edges[2 * e + 1];
        if (only_inside && !(v_valid[a] && v_valid[b]))
            continue;
        const int32_t u = old_to_new[a], w = old_to_new[b];
        if (u == w)
            continue;
        eset.push_back(edge_key(u, w));
        keep_edge[e] = 1;
    }
    int64_t fo = 0;
    if (nf > 0)
        new_face_off[0] = 0;
    int32_t kept = 0;
    std::vector<int32_t> fv;
    for (int32_t f = 0; f < nf; ++f)
    {
        if (only_inside && !face_valid(face_off, face_vts, v_valid, f))
        {
            remove_face[f] = 1;
            continue;
        }
        fv.clear();
        for (int64_t j = face_off[f]; j < face_off[f + 1]; ++j)
        {
            const int32_t c = old_to_new[face_vts[j]];
            if (std::find(fv.begin(), fv.end(), c) == fv.end())
                fv.push_back(c);
        }
        if (fv.size() < 2)
            remove_face[f] = 1;
        else if (fv.size() == 2)
        {
            eset.push_back(edge_key(fv[0], fv[1]));
            remove_face[f] = 1;
        }
        else
        {
            remove_face[f] = 0;
            for (int32_t c : fv)
                new_face_vts[fo++] = c;
            new_face_off[++kept] = fo;
        }
    }
    std::sort(eset.begin(), eset.end());
    eset.erase(std::unique(eset.begin(), eset.end()), eset.end());
    *n_new_edges = (int64_t)eset.size();
    if (new_edges)
    {
        if (edge_cap < *n_new_edges)
            return VC_ERR_CAPACITY;
        for (size_t i = 0; i < eset.size(); ++i)
        {
            new_edges[2 * i] = (int32_t)(eset[i] >> 32);
            new_edges[2 * i + 1] = (int32_t)(eset[i] & 0xFFFFFFFFull);
        }
    }
    return VC_OK;
}

int vc_voro_part_inside(int32_t nv, int32_t nf, const int64_t* face_off,
                        const int32_t* face_vts, const uint8_t* v_valid,
                        int32_t* vts_idx, int64_t vts_cap, int64_t* n_vts,
                        int32_t* edges_out, int64_t edge_cap, int64_t* n_edges,
                        uint32_t* tris, int64_t tri_cap, int64_t* n_tris)
{
    int rc = check_complex(nv, 0, nf, nullptr, face_off, face_vts);
    if (rc != VC_OK || (nv > 0 && !v_valid) || !n_vts || !n_edges || !n_tris)
        return VC_ERR_INVALID;
    std::vector<int32_t> vid, new_id((size_t)nv, 0);
    for (int32_t v = 0; v < nv; ++v)
        if (v_valid[v])
        {
            new_id[(size_t)v] = (int32_t)vid.size();
            vid.push_back(v);
        }
    // triangles of every face whose three corners are valid (the face itself
    // need not be), renumbered; their sides, deduplicated and sorted
    std::vector<uint32_t> t_out;
    std::vector<uint64_t> eset;
    std::vector<int32_t> ft;
    for (int32_t f = 0; f < nf; ++f)
    {
        zigzag(face_vts + face_off[f], (int)(face_off[f + 1] - face_off[f]), ft);
        for (size_t t = 0; t < ft.size(); t += 3)
        {
            if (!(v_valid[ft[t]] && v_valid[ft[t + 1]] && v_valid[ft[t + 2]]))
                continue;
            int32_t c[3];
            for (int k = 0; k < 3; ++k)
            {
                c[k] = new_id[(size_t)ft[t + k]];
                t_out.push_back((uint32_t)c[k]);
            }
            for (int k = 0; k < 3; ++k)
                eset.push_back(edge_key(c[k], c[(k + 1) % 3]));
        }
    }
    std::sort(eset.begin(), eset.end());
    eset.erase(std::unique(eset.begin(), eset.end()), eset.end());
    *n_vts = (int64_t)vid.size();
    *n_edges = (int64_t)eset.size();
    *n_tris = (int64_t)(t_out.size() / 3);
    if ((vts_idx && vts_cap < *n_vts) || (edges_out && edge_cap < *n_edges) ||
        (tris && tri_cap < *n_tris))
        return VC_ERR_CAPACITY;
    if (vts_idx)
        std::copy(vid.begin(), vid.end(), vts_idx);
    if (edges_out)
        for (size_t i = 0; i < eset.size(); ++i)
        {
            edges_out[2 * i] = (int32_t)(eset[i] >> 32);
            edges_out[2 * i + 1] = (int32_t)(eset[i] & 0xFFFFFFFFull);
        }
    if (tris)
        std::copy(t_out.begin(), t_out.end(), tris);
    return VC_OK;
}

} // extern "C"
