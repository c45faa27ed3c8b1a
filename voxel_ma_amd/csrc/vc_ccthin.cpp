// vc_ccthin.cpp -- host-side cell-complex finalize and FIFO simple-pair
// collapse behind the C ABI (vc_cc_finalize, vc_cc_prune).
//
// The collapse order is a FIFO whose pushes depend on the reference counts at
// the moment of each pop (src/ccthin.cpp:534-769), so the result (which edge
// of a face with two free edges survives, which end of a path keeps its
// vertex) is a property of that exact sequence; it runs sequentially here.
// Adjacency is stored as CSR arrays: vertex -> edges in edge order, edge ->
// faces in face order (src/cellcomplex.cpp:456-474).
#include <cstdint>
#include <cstring>
#include <deque>
#include <unordered_map>
#include <vector>

#include "../../include/vc_api.h"

namespace
{

struct PairHash
{
    size_t operator()(uint64_t k) const
    {
        k ^= k >> 33;
        k *= 0xff51afd7ed558ccdull;
        k ^= k >> 33;
        return (size_t)k;
    }
};

inline uint64_t key(int32_t a, int32_t b)
{
    return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}

struct Csr
{
    std::vector<int64_t> off;
    std::vector<int32_t> idx;
};

} // namespace

extern "C" int vc_cc_finalize(int32_t nv, int32_t ne_in, const int32_t* edges_in,
                              int32_t nf, const int64_t* face_off,
                              const int32_t* face_vts, int32_t* edges_out,
                              int32_t edge_cap, int32_t* ne_out,
                              int32_t* face_edges_out)
{
    if (nv < 0 || ne_in < 0 || nf < 0 || !ne_out ||
        (ne_in && !edges_in) || (nf && (!face_off || !face_vts || !face_edges_out)))
        return VC_ERR_INVALID;
    std::unordered_map<uint64_t, int32_t, PairHash> index;
    index.reserve((size_t)ne_in + (nf ? (size_t)face_off[nf] : 0));
    std::vector<int32_t> edges;
    edges.reserve(2 * (size_t)ne_in);
    for (int32_t e = 0; e < ne_in; ++e)
    {
        const int32_t a = edges_in[2 * e], b = edges_in[2 * e + 1];
        if (a < 0 || b < 0 || a >= nv || b >= nv)
            return VC_ERR_INVALID;
        auto ins = index.emplace(key(a, b), (int32_t)(edges.size() / 2));
        if (ins.second)
        {
            edges.push_back(a);
            edges.push_back(b);
        }
    }
    for (int32_t f = 0; f < nf; ++f)
    {
        const int64_t n = face_off[f + 1] - face_off[f];
        if (n < 0)
            return VC_ERR_INVALID;
        for (int64_t i = 0; i < n; ++i)
        {
            const int32_t u = face_vts[face_off[f] + i];
            const int32_t w = face_vts[face_off[f] + (i + 1) % n];
            if (u < 0 || w < 0 || u >= nv || w >= nv)
                return VC_ERR_INVALID;
            const int32_t a = u < w ? u : w, b = u < w ? w : u;
            auto ins = index.emplace(key(a, b), (int32_t)(edges.size() / 2));
            if (ins.second)
            {
                edges.push_back(a);
                edges.push_back(b);
            }
            face_edges_out[face_off[f] + i] = ins.first->second;
        }
    }
    const int32_t ne = (int32_t)(edges.size() / 2);
    *ne_out = ne;
    if (ne > edge_cap || (ne && !edges_out))
        return VC_ERR_CAPACITY;
    std::memcpy(edges_out, edges.data(), sizeof(int32_t) * edges.size());
    return VC_OK;
}

extern "C" int vc_cc_prune(const vc_cc_view* cc, const float* em,
                           const float* fm, float f_t, float l_t, uint8_t* rm_v,
                           uint8_t* rm_e, uint8_t* rm_f, int32_t* ref_e)
{
    if (!cc || cc->nv < 0 || cc->ne < 0 || cc->nf < 0)
        return VC_ERR_INVALID;
    const int32_t nv = cc->nv, ne = cc->ne, nf = cc->nf;
    if ((ne && (!cc->edges || !em || !rm_e || !ref_e)) ||
        (nf && (!cc->face_off || !cc->face_edges || !fm || !rm_f)) ||
        (nv && !rm_v))
        return VC_ERR_INVALID;
    const int32_t* E = cc->edges;
    const int64_t* fo = cc->face_off;
    const int32_t* fe = cc->face_edges;
    // CSR adjacency
    Csr ve, ef;
    ve.off.assign((size_t)nv + 1, 0);
    ef.off.assign((size_t)ne + 1, 0);
    for (int32_t e = 0; e < ne; ++e)
    {
        if (E[2 * e] < 0 || E[2 * e] >= nv || E[2 * e + 1] < 0 || E[2 * e + 1] >= nv)
            return VC_ERR_INVALID;
        ve.off[E[2 * e] + 1]++;
        ve.off[E[2 * e + 1] + 1]++;
    }
    const int64_t nfe = nf ? fo[nf] : 0;
    for (int64_t i = 0; i < nfe; ++i)
    {
        if (fe[i] < 0 || fe[i] >= ne)
            return VC_ERR_INVALID;
        ef.off[fe[i] + 1]++;
    }
    for (int32_t v = 0; v < nv; ++v)
        ve.off[v + 1] += ve.off[v];
    for (int32_t e = 0; e < ne; ++e)
        ef.off[e + 1] += ef.off[e];
    ve.idx.resize((size_t)ve.off[nv]);
    ef.idx.resize((size_t)ef.off[ne]);
    {
        std::vector<int64_t> fill(ve.off.begin(), ve.off.end() - 1);
        for (int32_t e = 0; e < ne; ++e)
        {
            ve.idx[fill[E[2 * e]]++] = e;
            ve.idx[fill[E[2 * e + 1]]++] = e;
        }
        std::vector<int64_t> fill2(ef.off.begin(), ef.off.end() - 1);
        for (int32_t f = 0; f < nf; ++f)
            for (int64_t i = fo[f]; i < fo[f + 1]; ++i)
                ef.idx[fill2[fe[i]]++] = f;
    }
    std::vector<int32_t> rv((size_t)nv);
    for (int32_t v = 0; v < nv; ++v)
        rv[v] = (int32_t)(ve.off[v + 1] - ve.off[v]);
    for (int32_t e = 0; e < ne; ++e)
        ref_e[e] = (int32_t)(ef.off[e + 1] - ef.off[e]);
    if (nv)
        std::memset(rm_v, 0, (size_t)nv);
    if (ne)
        std::memset(rm_e, 0, (size_t)ne);
    if (nf)
        std::memset(rm_f, 0, (size_t)nf);

    // simple pairs: face-edge (face, edge) or edge-vertex (edge, vertex)
    struct SPair
    {
        bool fe;
        int32_t cell, face_of;
    };
    std::deque<SPair> q;
    for (int32_t e = 0; e < ne; ++e)
        if (ref_e[e] == 1)
        {
            const int32_t f = ef.idx[ef.off[e]];
            if (fm[f] < f_t)
                q.push_back({true, f, e});
        }
    for (int32_t v = 0; v < nv; ++v)
        if (rv[v] == 1)
        {
            const int32_t e = ve.idx[ve.off[v]];
            if (em[e] < l_t)
                q.push_back({false, e, v});
        }
    // first non-removed edge of v / face of e, in adjacency order
    auto first_edge = [&](int32_t v) -> int32_t {
        for (int64_t i = ve.off[v]; i < ve.off[v + 1]; ++i)
            if (!rm_e[ve.idx[i]])
                return ve.idx[i];
        return -1;
    };
    auto first_face = [&](int32_t e) -> int32_t {
        for (int64_t i = ef.off[e]; i < ef.off[e + 1]; ++i)
            if (!rm_f[ef.idx[i]])
                return ef.idx[i];
        return -1;
    };
    while (!q.empty())
    {
        const SPair p = q.front();
        q.pop_front();
        if (p.fe ? rm_f[p.cell] : rm_e[p.cell])
            continue;
        if (p.fe ? ref_e[p.face_of] != 1 : rv[p.face_of] != 1)
            return VC_ERR_STATE; // the reference's "logic err" exit
        if (!p.fe)
        {
            const int32_t e = p.cell, v = p.face_of;
            rm_e[e] = 1;
            rm_v[v] = 1;
            const int32_t a = E[2 * e], b = E[2 * e + 1];
            rv[a]--;
            rv[b]--;
            const int32_t other = a != v ? a : b;
            if (rv[other] == 1)
            {
                const int32_t e2 = first_edge(other);
                if (e2 >= 0 && em[e2] < l_t)
                    q.push_back({false, e2, other});
            }
        }
        else
        {
            const int32_t f = p.cell, e = p.face_of;
            rm_f[f] = 1;
            rm_e[e] = 1;
            ref_e[e]--;
            for (int64_t i = fo[f]; i < fo[f + 1]; ++i)
            {
                const int32_t e2 = fe[i];
                if (rm_e[e2])
                    continue;
                if (--ref_e[e2] == 1)
                {
                    const int32_t f2 = first_face(e2);
                    if (f2 >= 0 && fm[f2] < f_t)
                        q.push_back({true, f2, e2});
                }
            }
            for (int k = 0; k < 2; ++k)
            {
                const int32_t vi = E[2 * e + k];
                if (--rv[vi] == 1)
                {
                    const int32_t e2 = first_edge(vi);
                    if (e2 >= 0 && em[e2] < l_t)
                        q.push_back({false, e2, vi});
                }
            }
        }
    }
    return VC_OK;
}
