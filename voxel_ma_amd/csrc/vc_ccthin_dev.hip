// vc_ccthin_dev.hip -- the FIFO simple-pair collapse of vc_cc_prune on the
// device, with the same result (src/ccthin.cpp:201-424, :514-771).
//
// The reference pops one simple pair at a time; a pop removes its cells,
// decrements reference counts and pushes the pairs that became simple, so
// the queue is a sequence of generations: generation g+1 is what the pops
// of generation g pushed, in push order.  One generation is processed in
// parallel and reproduces the sequential pops exactly:
//   * a pair takes effect iff its cell is still there at its turn.  Faces
//     are removed only by face-edge pairs, so a face-edge pair takes effect
//     iff it is the first pair of the generation on its face (and the face
//     was not removed before); edges are removed by the effective face-edge
//     pairs on them and by edge-vertex pairs, so an edge-vertex pair takes
//     effect iff it is the first of those on its edge;
//   * every decrement of an element's count is an event with the key (index
//     of the popping pair, step within the pop): a face-edge pop decrements
//     the face's edges at steps 0..|f|-1 and then the edge's two ends at
//     steps |f|, |f|+1; an edge-vertex pop decrements both ends at step 0;
//   * an element pushes a pair at the event that brings its count to 1,
//     which is its (r0 - 1)-th event in key order (r0 = count at the start
//     of the generation).  The pushed pair names the element's one cell left
//     at that moment: the first in adjacency order not removed before the
//     generation nor by an event with a smaller or equal key;
//   * pushes are written to slots numbered by their keys (per pair |f| + 2
//     slots for a face-edge pop, 1 for an edge-vertex pop), and compacting
//     the slots gives the next generation in push order.
// The reference exits with "logic err" when a pair it takes has a count
// other than 1; those cases return VC_ERR_STATE here, as in vc_cc_prune.
// An effective face-edge pair whose edge an edge-vertex pair of the same
// generation removed cannot occur in a complex whose faces are polygons
// (an edge-vertex pair needs its vertex to have one edge, a face has two
// at each of its vertices); it is reported as VC_ERR_STATE too.
//
// Adjacency (vertex -> edges in edge order, edge -> faces in face order) is
// built on the device: counts, a scan, an atomic fill and a sort of every
// (short) list.
#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "vc_internal.h"
#include "../../include/vc_cc.h"

namespace
{

constexpr int kB = 256;
constexpr int kScanB = 1024;
constexpr int kErrLogic = 1;   // a taken pair with a count other than 1
constexpr int kErrTwice = 2;   // an edge removed twice in one generation
constexpr int kErrGuard = 4;   // a union-find iteration guard tripped
// adjacency build flags (bits): an id out of range (nothing can be built);
// an edge with equal ends or a face listing a side twice (the adjacency is
// built, the collapse refuses the complex)
constexpr int kAdjRange = 16, kAdjDegenerate = 32;

unsigned nblk(int64_t n) { return (unsigned)((n + kB - 1) / kB); }

// ------------------------------------------------------------------ scans
// exclusive scan of up to kScanB * kScanB uint32 in place: per-chunk scan,
// a one-block scan of the chunk totals, then the chunk offsets added
__global__ void __launch_bounds__(kScanB)
    scan_chunk(uint32_t* a, int64_t n, uint32_t* sums)
{
    __shared__ uint32_t sh[kScanB / 64];
    const int64_t i = (int64_t)blockIdx.x * kScanB + threadIdx.x;
    const uint32_t v = i < n ? a[i] : 0u;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1)
    {
        const uint32_t t = __shfl_up(x, d);
        if (lane >= d)
            x += t;
    }
    if (lane == 63)
        sh[wv] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int w = 0; w < kScanB / 64; ++w)
    {
        off += w < wv ? sh[w] : 0u;
        tot += sh[w];
    }
    if (i < n)
        a[i] = off + x - v;
    if (threadIdx.x == 0 && sums)
        sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanB)
    scan_add(uint32_t* a, int64_t n, const uint32_t* sums)
{
    const int64_t i = (int64_t)blockIdx.x * kScanB + threadIdx.x;
    if (i < n)
        a[i] += sums[blockIdx.x];
}

// ------------------------------------------------------------- adjacency
__global__ void __launch_bounds__(kB)
    deg_v(const int32_t* __restrict__ E, int32_t ne, int32_t nv,
          uint32_t* __restrict__ cnt, int* err)
{
    const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (e >= ne)
        return;
    const int32_t a = E[2 * e], b = E[2 * e + 1];
    if (a < 0 || b < 0 || a >= nv || b >= nv)
    {
        atomicOr(err, kAdjRange);
        return;
    }
    if (a == b)
        atomicOr(err, kAdjDegenerate);
    atomicAdd(&cnt[a], 1u);
    atomicAdd(&cnt[b], 1u);
}

// per face: side ids in range (else nothing is counted), sides listed
// twice flagged; edge face counts
__global__ void __launch_bounds__(kB)
    deg_e(const int64_t* __restrict__ fo, const int32_t* __restrict__ fe,
          int32_t nf, int32_t ne, uint32_t* __restrict__ cnt, int* err)
{
    const int64_t f = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (f >= nf)
        return;
    for (int64_t i = fo[f]; i < fo[f + 1]; ++i)
    {
        const int32_t e = fe[i];
        if (e < 0 || e >= ne)
        {
            atomicOr(err, kAdjRange);
            return;
        }
        bool twice = false;
        for (int64_t k = fo[f]; k < i && !twice; ++k)
            twice = fe[k] == e;
        if (twice)
            atomicOr(err, kAdjDegenerate);
    }
    for (int64_t i = fo[f]; i < fo[f + 1]; ++i)
        atomicAdd(&cnt[fe[i]], 1u);
}

__global__ void __launch_bounds__(kB)
    fill_ve(const int32_t* __restrict__ E, int32_t ne,
            const uint32_t* __restrict__ off, uint32_t* __restrict__ fill,
            int32_t* __restrict__ ve)
{
    const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (e >= ne)
        return;
    for (int k = 0; k < 2; ++k)
    {
        const int32_t v = E[2 * e + k];
        ve[off[v] + atomicAdd(&fill[v], 1u)] = (int32_t)e;
    }
}

__global__ void __launch_bounds__(kB)
    fill_ef(const int64_t* __restrict__ fo, const int32_t* __restrict__ fe,
            int32_t nf, const uint32_t* __restrict__ off,
            uint32_t* __restrict__ fill, int32_t* __restrict__ ef)
{
    const int64_t f = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (f >= nf)
        return;
    for (int64_t i = fo[f]; i < fo[f + 1]; ++i)
    {
        const int32_t e = fe[i];
        ef[off[e] + atomicAdd(&fill[e], 1u)] = (int32_t)f;
    }
}

// ascending ids in every list (insertion sort: lists are short)
__global__ void __launch_bounds__(kB)
    sort_lists(const uint32_t* __restrict__ off, int32_t n,
               int32_t* __restrict__ idx)
{
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t lo = off[i], hi = off[i + 1];
    for (uint32_t k = lo + 1; k < hi; ++k)
    {
        const int32_t x = idx[k];
        uint32_t m = k;
        while (m > lo && idx[m - 1] > x)
        {
            idx[m] = idx[m - 1];
            --m;
        }
        idx[m] = x;
    }
}

// the start of a collapse: every count is its adjacency list's length
// (src/ccthin.cpp:236-242), nothing removed
__global__ void __launch_bounds__(kB)
    reset_state(int32_t nv, int32_t ne, int32_t nf,
                const uint32_t* __restrict__ ve_off,
                const uint32_t* __restrict__ ef_off, int32_t* __restrict__ ref_v,
                int32_t* __restrict__ ref_e, int32_t* __restrict__ rg_f,
                int32_t* __restrict__ rg_e, int32_t* __restrict__ first_f,
                int32_t* __restrict__ first_e, uint8_t* __restrict__ rm_v)
{
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i < nv)
    {
        ref_v[i] = (int32_t)(ve_off[i + 1] - ve_off[i]);
        rm_v[i] = 0;
    }
    if (i < ne)
    {
        ref_e[i] = (int32_t)(ef_off[i + 1] - ef_off[i]);
        rg_e[i] = -1;
        first_e[i] = INT_MAX;
    }
    if (i < nf)
    {
        rg_f[i] = -1;
        first_f[i] = INT_MAX;
    }
}

// initial queue flags: edges (face-edge pairs) first, then vertices
__global__ void __launch_bounds__(kB)
    init_flags(int32_t ne, int32_t nv, const int32_t* __restrict__ ref_e,
               const int32_t* __restrict__ ref_v,
               const uint32_t* __restrict__ ef_off, const int32_t* __restrict__ ef,
               const uint32_t* __restrict__ ve_off, const int32_t* __restrict__ ve,
               const float* __restrict__ em, const float* __restrict__ fm,
               float f_t, float l_t, uint32_t* __restrict__ flag)
{
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i >= (int64_t)ne + nv)
        return;
    bool take;
    if (i < ne)
        take = ref_e[i] == 1 && fm[ef[ef_off[i]]] < f_t;
    else
    {
        const int64_t v = i - ne;
        take = ref_v[v] == 1 && em[ve[ve_off[v]]] < l_t;
    }
    flag[i] = take ? 1u : 0u;
}

// pair: t = 1 face-edge (c = face, o = edge), t = 0 edge-vertex (c = edge,
// o = vertex)
struct Pairs
{
    uint8_t* t;
    int32_t* c;
    int32_t* o;
};

__global__ void __launch_bounds__(kB)
    init_pairs(int32_t ne, int32_t nv, const uint32_t* __restrict__ pos,
               const uint32_t* __restrict__ flag_end,
               const uint32_t* __restrict__ ef_off, const int32_t* __restrict__ ef,
               const uint32_t* __restrict__ ve_off, const int32_t* __restrict__ ve,
               Pairs P)
{
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    const int64_t n = (int64_t)ne + nv;
    if (i >= n)
        return;
    const uint32_t next = i + 1 < n ? pos[i + 1] : *flag_end;
    if (next == pos[i])
        return;
    const uint32_t k = pos[i];
    if (i < ne)
    {
        P.t[k] = 1;
        P.c[k] = ef[ef_off[i]];
        P.o[k] = (int32_t)i;
    }
    else
    {
        const int32_t v = (int32_t)(i - ne);
        P.t[k] = 0;
        P.c[k] = ve[ve_off[v]];
        P.o[k] = v;
    }
}

// ------------------------------------------------------------ generation
struct Cx
{
    const int32_t* E;
    const int64_t* fo;
    const int32_t* fe;
    const uint32_t* ve_off;
    const int32_t* ve;
    const uint32_t* ef_off;
    const int32_t* ef;
    const float* em;
    const float* fm;
    int32_t* ref_e;
    int32_t* ref_v;
    int32_t* rg_f; // generation that removed the face, -1 = present
    int32_t* rj_f; // index of the removing pair in that generation
    int32_t* rg_e;
    int32_t* rj_e;
    uint8_t* rm_v;
    int32_t* first_f;
    int32_t* first_e;
    int* err;
};

__global__ void __launch_bounds__(kB) first_face(Cx C, Pairs P, int32_t np)
{
    const int32_t j = blockIdx.x * kB + threadIdx.x;
    if (j >= np || !P.t[j])
        return;
    const int32_t f = P.c[j];
    if (C.rg_f[f] < 0)
        atomicMin(&C.first_f[f], j);
}

// face-edge pairs: effective iff first on their face; the effective ones and
// the edge-vertex pairs compete for their edge; slot counts
__global__ void __launch_bounds__(kB)
    first_edge(Cx C, Pairs P, int32_t np, uint8_t* __restrict__ eff,
               uint32_t* __restrict__ slots)
{
    const int32_t j = blockIdx.x * kB + threadIdx.x;
    if (j >= np)
        return;
    if (P.t[j])
    {
        const int32_t f = P.c[j];
        const bool e1 = C.rg_f[f] < 0 && C.first_f[f] == j;
        eff[j] = e1 ? 1 : 0;
        if (e1)
            atomicMin(&C.first_e[P.o[j]], j);
        slots[j] = (uint32_t)(C.fo[f + 1] - C.fo[f]) + 2u;
    }
    else
    {
        const int32_t e = P.c[j];
        if (C.rg_e[e] < 0)
            atomicMin(&C.first_e[e], j);
        slots[j] = 1u;
    }
}

__global__ void __launch_bounds__(kB)
    decide(Cx C, Pairs P, int32_t np, uint8_t* __restrict__ eff)
{
    const int32_t j = blockIdx.x * kB + threadIdx.x;
    if (j >= np)
        return;
    if (P.t[j])
    {
        if (!eff[j])
            return;
        const int32_t e = P.o[j];
        if (C.first_e[e] != j)
            atomicMax(C.err, kErrTwice);
        if (C.ref_e[e] != 1)
            atomicMax(C.err, kErrLogic);
    }
    else
    {
        const int32_t e = P.c[j];
        const bool e1 = C.rg_e[e] < 0 && C.first_e[e] == j;
        eff[j] = e1 ? 1 : 0;
        if (e1 && C.ref_v[P.o[j]] != 1)
            atomicMax(C.err, kErrLogic);
    }
}

__global__ void __launch_bounds__(kB)
    apply(Cx C, Pairs P, int32_t np, const uint8_t* __restrict__ eff, int32_t g)
{
    const int32_t j = blockIdx.x * kB + threadIdx.x;
    if (j >= np)
        return;
    if (P.t[j])
    {
        const int32_t f = P.c[j], e = P.o[j];
        C.first_f[f] = INT_MAX;
        C.first_e[e] = INT_MAX;
        if (!eff[j])
            return;
        C.rg_f[f] = g;
        C.rj_f[f] = j;
        C.rg_e[e] = g;
        C.rj_e[e] = j;
        for (int64_t i = C.fo[f]; i < C.fo[f + 1]; ++i)
            atomicSub(&C.ref_e[C.fe[i]], 1);
        atomicSub(&C.ref_v[C.E[2 * e]], 1);
        atomicSub(&C.ref_v[C.E[2 * e + 1]], 1);
    }
    else
    {
        const int32_t e = P.c[j], v = P.o[j];
        C.first_e[e] = INT_MAX;
        if (!eff[j])
            return;
        C.rg_e[e] = g;
        C.rj_e[e] = j;
        C.rm_v[v] = 1;
        atomicSub(&C.ref_v[C.E[2 * e]], 1);
        atomicSub(&C.ref_v[C.E[2 * e + 1]], 1);
    }
}

__device__ __forceinline__ int64_t mkkey(int32_t j, int64_t s)
{
    return ((int64_t)j << 32) | s;
}

// position of edge x among the sides of face f
__device__ __forceinline__ int64_t side_pos(const Cx& C, int32_t f, int32_t x)
{
    for (int64_t i = C.fo[f]; i < C.fo[f + 1]; ++i)
        if (C.fe[i] == x)
            return i - C.fo[f];
    return 0;
}

// key of the decrement of vertex x by the removal of edge e in generation g
__device__ __forceinline__ int64_t vert_event(const Cx& C, const Pairs& P,
                                              int32_t e, int32_t x)
{
    const int32_t j = C.rj_e[e];
    if (!P.t[j])
        return mkkey(j, 0);
    const int32_t f = P.c[j];
    return mkkey(j, (C.fo[f + 1] - C.fo[f]) + (C.E[2 * e] == x ? 0 : 1));
}

// the event (j, s) of edge x: if it brings x's count to 1, push the
// face-edge pair of the face left at that moment
__device__ void edge_event(const Cx& C, const Pairs& P, int32_t g, int32_t x,
                           int64_t K, float f_t, uint32_t slot,
                           uint32_t* sflag, Pairs N)
{
    int n = 0, rank = 0;
    for (uint32_t i = C.ef_off[x]; i < C.ef_off[x + 1]; ++i)
    {
        const int32_t f = C.ef[i];
        if (C.rg_f[f] != g)
            continue;
        ++n;
        rank += mkkey(C.rj_f[f], side_pos(C, f, x)) < K ? 1 : 0;
    }
    const int r0 = C.ref_e[x] + n;
    if (r0 < 2 || rank != r0 - 2)
        return;
    for (uint32_t i = C.ef_off[x]; i < C.ef_off[x + 1]; ++i)
    {
        const int32_t f = C.ef[i];
        const int32_t rg = C.rg_f[f];
        if (rg >= 0 && (rg != g || mkkey(C.rj_f[f], side_pos(C, f, x)) <= K))
            continue;
        if (C.fm[f] < f_t)
        {
            N.t[slot] = 1;
            N.c[slot] = f;
            N.o[slot] = x;
            sflag[slot] = 1u;
        }
        return;
    }
}

__device__ void vert_event_push(const Cx& C, const Pairs& P, int32_t g,
                                int32_t x, int64_t K, float l_t, uint32_t slot,
                                uint32_t* sflag, Pairs N)
{
    int n = 0, rank = 0;
    for (uint32_t i = C.ve_off[x]; i < C.ve_off[x + 1]; ++i)
    {
        const int32_t e = C.ve[i];
        if (C.rg_e[e] != g)
            continue;
        ++n;
        rank += vert_event(C, P, e, x) < K ? 1 : 0;
    }
    const int r0 = C.ref_v[x] + n;
    if (r0 < 2 || rank != r0 - 2)
        return;
    for (uint32_t i = C.ve_off[x]; i < C.ve_off[x + 1]; ++i)
    {
        const int32_t e = C.ve[i];
        const int32_t rg = C.rg_e[e];
        if (rg >= 0 && (rg != g || vert_event(C, P, e, x) <= K))
            continue;
        if (C.em[e] < l_t)
        {
            N.t[slot] = 0;
            N.c[slot] = e;
            N.o[slot] = x;
            sflag[slot] = 1u;
        }
        return;
    }
}

__global__ void __launch_bounds__(kB)
    trigger(Cx C, Pairs P, int32_t np, const uint8_t* __restrict__ eff,
            int32_t g, const uint32_t* __restrict__ slot_off, float f_t,
            float l_t, uint32_t* __restrict__ sflag, Pairs N)
{
    const int32_t j = blockIdx.x * kB + threadIdx.x;
    if (j >= np || !eff[j])
        return;
    const uint32_t s0 = slot_off[j];
    if (P.t[j])
    {
        const int32_t f = P.c[j], e = P.o[j];
        const int64_t L = C.fo[f + 1] - C.fo[f];
        for (int64_t s = 0; s < L; ++s)
        {
            const int32_t x = C.fe[C.fo[f] + s];
            if (x != e)
                edge_event(C, P, g, x, mkkey(j, s), f_t, s0 + (uint32_t)s,
                           sflag, N);
        }
        for (int k = 0; k < 2; ++k)
            vert_event_push(C, P, g, C.E[2 * e + k], mkkey(j, L + k), l_t,
                            s0 + (uint32_t)(L + k), sflag, N);
    }
    else
    {
        const int32_t e = P.c[j], v = P.o[j];
        const int32_t a = C.E[2 * e], b = C.E[2 * e + 1];
        vert_event_push(C, P, g, a != v ? a : b, mkkey(j, 0), l_t, s0, sflag,
                        N);
    }
}

__global__ void __launch_bounds__(kB)
    compact(const uint32_t* __restrict__ pos, uint32_t ns,
            const uint32_t* __restrict__ total, Pairs S, Pairs N)
{
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i >= ns)
        return;
    const uint32_t next = i + 1 < ns ? pos[i + 1] : *total;
    if (next == pos[i])
        return;
    const uint32_t k = pos[i];
    N.t[k] = S.t[i];
    N.c[k] = S.c[i];
    N.o[k] = S.o[i];
}

__global__ void __launch_bounds__(kB) fill_u32(uint32_t* a, int64_t n, uint32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i < n)
        a[i] = v;
}

__global__ void __launch_bounds__(kB)
    results(int32_t ne, int32_t nf, const int32_t* __restrict__ rg_e,
            const int32_t* __restrict__ rg_f, uint8_t* __restrict__ rm_e,
            uint8_t* __restrict__ rm_f)
{
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i < ne)
        rm_e[i] = rg_e[i] >= 0 ? 1 : 0;
    if (i < nf)
        rm_f[i] = rg_f[i] >= 0 ? 1 : 0;
}

// ----------------------------------------------- remaining complex (R12)
// flags, scanned in place: kept vertices; kept isolated edges (not removed,
// no face left on them, src/ccthin.cpp:163); triangles per kept face
// (util::simpleTriangulate gives |f| - 2 for |f| >= 3, none otherwise)
__global__ void __launch_bounds__(kB)
    rem_flags(int32_t nv, int32_t ne, int32_t nf, const uint8_t* __restrict__ rm_v,
              const uint8_t* __restrict__ rm_e, const int32_t* __restrict__ ref_e,
              const uint8_t* __restrict__ rm_f, const int64_t* __restrict__ fo,
              uint32_t* __restrict__ vflag, uint32_t* __restrict__ eflag,
              uint32_t* __restrict__ tcnt)
{
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i < nv)
        vflag[i] = rm_v[i] ? 0u : 1u;
    if (i < ne)
        eflag[i] = !rm_e[i] && ref_e[i] == 0 ? 1u : 0u;
    if (i < nf)
    {
        const int64_t n = fo[i + 1] - fo[i];
        tcnt[i] = !rm_f[i] && n >= 3 ? (uint32_t)(n - 2) : 0u;
    }
}

// util::compactify (src/geomalgo.cpp:215-235): a kept vertex's new id is its
// rank among the kept ones; an id that is not kept maps to 0, as the
// reference's unordered_map::operator[] does
__device__ __forceinline__ int32_t rem_map(const uint32_t* vpos, int32_t v)
{
    return vpos[v + 1] != vpos[v] ? (int32_t)vpos[v] : 0;
}

__global__ void __launch_bounds__(kB)
    rem_emit(int32_t nv, int32_t ne, int32_t nf, const int32_t* __restrict__ E,
             const int64_t* __restrict__ fo, const int32_t* __restrict__ fv,
             const uint32_t* __restrict__ vpos, const uint32_t* __restrict__ epos,
             const uint32_t* __restrict__ tpos, int32_t* __restrict__ vts_idx,
             int32_t* __restrict__ e_out, int32_t* __restrict__ t_out,
             int32_t* __restrict__ from_fi)
{
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i < nv && vts_idx && vpos[i + 1] != vpos[i])
        vts_idx[vpos[i]] = (int32_t)i;
    if (i < ne && e_out && epos[i + 1] != epos[i])
    {
        const int64_t k = epos[i];
        e_out[2 * k] = rem_map(vpos, E[2 * i]);
        e_out[2 * k + 1] = rem_map(vpos, E[2 * i + 1]);
    }
    if (i < nf && (t_out || from_fi) && tpos[i + 1] != tpos[i])
    {
        // zig-zag triangulation (src/geomalgo.cpp:13-35): (p[hi], p[lo],
        // p[lo+1]) then (p[hi], p[lo], p[hi-1]), alternating
        const int32_t* p = fv + fo[i];
        int hi = (int)(fo[i + 1] - fo[i]) - 1, lo = 0;
        bool inc = true;
        int64_t k = tpos[i];
        while (hi - 1 > lo)
        {
            const int32_t c = inc ? p[lo + 1] : p[hi - 1];
            if (t_out)
            {
                t_out[3 * k] = rem_map(vpos, p[hi]);
                t_out[3 * k + 1] = rem_map(vpos, p[lo]);
                t_out[3 * k + 2] = rem_map(vpos, c);
            }
            if (from_fi)
                from_fi[k] = (int32_t)i;
            if (inc)
                ++lo;
            else
                --hi;
            inc = !inc;
            ++k;
        }
    }
}

// getRemainedVts / getRemainedEdges / getRemainedFaces: kept ids in order
__global__ void __launch_bounds__(kB)
    kept_flags(int32_t nv, int32_t ne, int32_t nf, const uint8_t* __restrict__ rm_v,
               const uint8_t* __restrict__ rm_e, const int32_t* __restrict__ ref_e,
               const uint8_t* __restrict__ rm_f, uint32_t* __restrict__ vflag,
               uint32_t* __restrict__ eflag, uint32_t* __restrict__ fflag)
{
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i < nv)
        vflag[i] = rm_v[i] ? 0u : 1u;
    if (i < ne)
        eflag[i] = !rm_e[i] && ref_e[i] == 0 ? 1u : 0u;
    if (i < nf)
        fflag[i] = rm_f[i] ? 0u : 1u;
}

__global__ void __launch_bounds__(kB)
    kept_ids(int32_t nv, int32_t ne, int32_t nf, const uint32_t* __restrict__ vpos,
             const uint32_t* __restrict__ epos, const uint32_t* __restrict__ fpos,
             int32_t* __restrict__ v_ids, int32_t* __restrict__ e_ids,
             int32_t* __restrict__ f_ids)
{
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i < nv && v_ids && vpos[i + 1] != vpos[i])
        v_ids[vpos[i]] = (int32_t)i;
    if (i < ne && e_ids && epos[i + 1] != epos[i])
        e_ids[epos[i]] = (int32_t)i;
    if (i < nf && f_ids && fpos[i + 1] != fpos[i])
        f_ids[fpos[i]] = (int32_t)i;
}

// ----------------------------------------------- components (R14)
__device__ __forceinline__ int32_t uf_find(int32_t* par, int32_t i)
{
    int32_t p = par[i];
    while (p != i)
    {
        const int32_t g = par[p];
        if (g == p)
            break;
        // path halving (benign races: parents only ever decrease)
        par[i] = g;
        i = p;
        p = g;
    }
    return p;
}

// the smaller root wins; a failed CAS means another link succeeded, so the
// loop is lock-free; the bound is a safety net reported through err
__device__ void uf_union(int32_t* par, int32_t a, int32_t b, int* err)
{
    for (int guard = 0; guard < (1 << 20); ++guard)
    {
        a = uf_find(par, a);
        b = uf_find(par, b);
        if (a == b)
            return;
        if (a > b)
        {
            const int32_t t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(&par[b], b, a) == b)
            return;
    }
    atomicMax(err, kErrGuard);
}

__global__ void __launch_bounds__(kB) uf_init(int32_t* par, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i < n)
        par[i] = (int32_t)i;
}

// vertices joined by edges (2 ids per edge)
__global__ void __launch_bounds__(kB)
    uf_pairs(const int32_t* __restrict__ E, int64_t ne, int32_t* par, int* err)
{
    const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (e < ne)
        uf_union(par, E[2 * e], E[2 * e + 1], err);
}

// faces joined through the edges they share (edge -> faces lists)
__global__ void __launch_bounds__(kB)
    uf_faces(const uint32_t* __restrict__ ef_off, const int32_t* __restrict__ ef,
             int32_t ne, int32_t* par, int* err)
{
    const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (e >= ne)
        return;
    for (uint32_t k = ef_off[e] + 1; k < ef_off[e + 1]; ++k)
        uf_union(par, ef[ef_off[e]], ef[k], err);
}

// number of roots among par[0..n), block-reduced
__global__ void __launch_bounds__(kB)
    uf_roots(const int32_t* __restrict__ par, int64_t n,
             unsigned long long* __restrict__ count)
{
    __shared__ unsigned sh[kB / 64];
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    const bool r = i < n && par[i] == (int32_t)i;
    const unsigned long long b = __ballot(r);
    if ((threadIdx.x & 63) == 0)
        sh[threadIdx.x >> 6] = (unsigned)__popcll(b);
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned t = 0;
        for (int w = 0; w < kB / 64; ++w)
            t += sh[w];
        if (t)
            atomicAdd(count, (unsigned long long)t);
    }
}

// compute_conn_cmpnts(subset, 1): the vertices of the chosen faces, joined
// by the chosen faces' polygon sides
__global__ void __launch_bounds__(kB)
    subset_union(int32_t nf, const int64_t* __restrict__ fo,
                 const int32_t* __restrict__ fv, const uint8_t* __restrict__ face_in,
                 uint8_t* __restrict__ in_set, int32_t* par, int* err)
{
    const int64_t f = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (f >= nf || !face_in[f])
        return;
    const int64_t n = fo[f + 1] - fo[f];
    const int32_t* p = fv + fo[f];
    for (int64_t j = 0; j < n; ++j)
    {
        in_set[p[j]] = 1;
        uf_union(par, p[j], p[(j + 1) % n], err);
    }
}

__global__ void __launch_bounds__(kB)
    uf_roots_in(const int32_t* __restrict__ par, const uint8_t* __restrict__ in_set,
                int64_t n, unsigned long long* __restrict__ count)
{
    __shared__ unsigned sh[kB / 64];
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    const bool r = i < n && in_set[i] && par[i] == (int32_t)i;
    const unsigned long long b = __ballot(r);
    if ((threadIdx.x & 63) == 0)
        sh[threadIdx.x >> 6] = (unsigned)__popcll(b);
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned t = 0;
        for (int w = 0; w < kB / 64; ++w)
            t += sh[w];
        if (t)
            atomicAdd(count, (unsigned long long)t);
    }
}

// ----------------------------------------------- pair sets
constexpr uint64_t kEmpty = ~0ull;

__device__ __forceinline__ uint64_t key_of(int32_t a, int32_t b)
{
    return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}

__device__ __forceinline__ uint64_t mix(uint64_t k)
{
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    return k;
}

// open addressing; the table keeps the smallest value per key
__device__ void set_min(uint64_t* keys, unsigned long long* vals, uint64_t mask,
                        uint64_t key, unsigned long long v)
{
    uint64_t h = mix(key) & mask;
    for (;;)
    {
        const uint64_t prev = atomicCAS(reinterpret_cast<unsigned long long*>(&keys[h]),
                                        (unsigned long long)kEmpty,
                                        (unsigned long long)key);
        if (prev == kEmpty || prev == key)
        {
            atomicMin(&vals[h], v);
            return;
        }
        h = (h + 1) & mask;
    }
}

__device__ unsigned long long set_get(const uint64_t* keys,
                                      const unsigned long long* vals,
                                      uint64_t mask, uint64_t key)
{
    uint64_t h = mix(key) & mask;
    while (keys[h] != key)
        h = (h + 1) & mask;
    return vals[h];
}

__global__ void __launch_bounds__(kB)
    fill_u64(uint64_t* a, int64_t n, uint64_t v)
{
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i < n)
        a[i] = v;
}

// candidate pairs of util::computeEulerChar(vts, edges, tris): the given
// edges as they are, then each triangle's sides makeEdge-normalised
// (src/geomalgo.cpp:174-214); cand[2c], cand[2c+1] are the pair's ends
__global__ void __launch_bounds__(kB)
    euler_cands(const int32_t* __restrict__ E, int64_t ne,
                const int32_t* __restrict__ T, int64_t nt,
                int32_t* __restrict__ cand, uint64_t* keys,
                unsigned long long* vals, uint64_t mask)
{
    const int64_t c = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (c >= ne + 3 * nt)
        return;
    int32_t a, b;
    if (c < ne)
    {
        a = E[2 * c];
        b = E[2 * c + 1];
    }
    else
    {
        const int64_t t = (c - ne) / 3;
        const int k = (int)((c - ne) % 3);
        a = T[3 * t + k];
        b = T[3 * t + (k + 1) % 3];
        if (a > b)
        {
            const int32_t x = a;
            a = b;
            b = x;
        }
    }
    cand[2 * c] = a;
    cand[2 * c + 1] = b;
    set_min(keys, vals, mask, key_of(a, b), (unsigned long long)c);
}

// distinct pairs: candidates that are the first of their pair
__global__ void __launch_bounds__(kB)
    count_first(const int32_t* __restrict__ cand, int64_t nc, const uint64_t* keys,
                const unsigned long long* vals, uint64_t mask,
                unsigned long long* __restrict__ count)
{
    __shared__ unsigned sh[kB / 64];
    const int64_t c = (int64_t)blockIdx.x * kB + threadIdx.x;
    const bool first = c < nc && set_get(keys, vals, mask,
                                         key_of(cand[2 * c], cand[2 * c + 1])) ==
                                     (unsigned long long)c;
    const unsigned long long b = __ballot(first);
    if ((threadIdx.x & 63) == 0)
        sh[threadIdx.x >> 6] = (unsigned)__popcll(b);
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned t = 0;
        for (int w = 0; w < kB / 64; ++w)
            t += sh[w];
        if (t)
            atomicAdd(count, (unsigned long long)t);
    }
}

// triangles joined through shared (normalised) sides: the finalized
// remaining complex gives two triangles one edge id iff their sides are
// the same normalised pair
__global__ void __launch_bounds__(kB)
    tri_sides(const int32_t* __restrict__ T, int64_t nt, uint64_t* keys,
              unsigned long long* vals, uint64_t mask)
{
    const int64_t c = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (c >= 3 * nt)
        return;
    const int64_t t = c / 3;
    const int k = (int)(c % 3);
    int32_t a = T[3 * t + k], b = T[3 * t + (k + 1) % 3];
    if (a > b)
    {
        const int32_t x = a;
        a = b;
        b = x;
    }
    set_min(keys, vals, mask, key_of(a, b), (unsigned long long)t);
}

__global__ void __launch_bounds__(kB)
    tri_union(const int32_t* __restrict__ T, int64_t nt, const uint64_t* keys,
              const unsigned long long* vals, uint64_t mask, int32_t* par,
              int* err)
{
    const int64_t c = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (c >= 3 * nt)
        return;
    const int64_t t = c / 3;
    const int k = (int)(c % 3);
    int32_t a = T[3 * t + k], b = T[3 * t + (k + 1) % 3];
    if (a > b)
    {
        const int32_t x = a;
        a = b;
        b = x;
    }
    const int32_t first = (int32_t)set_get(keys, vals, mask, key_of(a, b));
    if (first != (int32_t)t)
        uf_union(par, (int32_t)t, first, err);
}

// device allocations released with their owner
struct Mem
{
    std::vector<void*> p;
    Mem() = default;
    Mem(const Mem&) = delete;
    Mem& operator=(const Mem&) = delete;
    ~Mem()
    {
        for (void* q : p)
            vc_free(q);
    }
    template <typename T>
    int get(vc_ctx* ctx, T** out, size_t count)
    {
        void* q = nullptr;
        const int rc = vc_alloc(ctx, &q, count * sizeof(T) + 16);
        if (rc == VC_OK)
            p.push_back(q);
        *out = static_cast<T*>(q);
        return rc;
    }
};

// exclusive scan of n <= kScanB^2 values; `sums` holds kScanB + 1 words
int scan_small(vc_ctx* ctx, uint32_t* a, int64_t n, uint32_t* sums)
{
    hipStream_t S = ctx->stream;
    const int64_t chunks = (n + kScanB - 1) / kScanB;
    if (chunks <= 1)
    {
        scan_chunk<<<1, kScanB, 0, S>>>(a, n, nullptr);
        VC_CHECK_LAUNCH(ctx);
        return VC_OK;
    }
    scan_chunk<<<(unsigned)chunks, kScanB, 0, S>>>(a, n, sums);
    VC_CHECK_LAUNCH(ctx);
    scan_chunk<<<1, kScanB, 0, S>>>(sums, chunks, nullptr);
    VC_CHECK_LAUNCH(ctx);
    scan_add<<<(unsigned)chunks, kScanB, 0, S>>>(a, n, sums);
    VC_CHECK_LAUNCH(ctx);
    return VC_OK;
}

int scan_any(vc_ctx* ctx, uint32_t* a, int64_t n, uint32_t* sums)
{
    if (n <= (int64_t)kScanB * kScanB)
        return scan_small(ctx, a, n, sums);
    return vc_scan_u32(ctx, a, n);
}

// hash table size: a power of two >= 2n + 1
uint64_t table_size(int64_t n)
{
    uint64_t H = 1024;
    while (H < 2 * (uint64_t)n + 1)
        H <<= 1;
    return H;
}

// number of roots of par[0..n) (waits)
int count_roots(vc_ctx* ctx, const int32_t* par, int64_t n,
                unsigned long long* dcount, int64_t* out)
{
    hipStream_t S = ctx->stream;
    VC_HIP(ctx, hipMemsetAsync(dcount, 0, sizeof(unsigned long long), S));
    if (n)
        uf_roots<<<nblk(n), kB, 0, S>>>(par, n, dcount);
    VC_CHECK_LAUNCH(ctx);
    unsigned long long h = 0;
    VC_HIP(ctx, hipMemcpyAsync(&h, dcount, sizeof(h), hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    *out = (int64_t)h;
    return VC_OK;
}

int read_err(vc_ctx* ctx, const int* err, int* out)
{
    VC_HIP(ctx, hipMemcpyAsync(out, err, sizeof(int), hipMemcpyDeviceToHost,
                               ctx->stream));
    VC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VC_OK;
}

} // namespace

// ------------------------------------------------------------ the handles
// a finalized complex resident on the device: the edge list, each face's
// vertices (polygon order) and side edge ids, and the adjacency (vertex ->
// edges in edge order, edge -> faces in face order)
struct vc_cc
{
    vc_ctx* ctx = nullptr;
    int32_t nv = 0, ne = 0, nf = 0;
    int64_t nfe = 0; // face sides
    int32_t* E = nullptr;
    int64_t* fo = nullptr;
    int32_t* fv = nullptr; // null when built from side ids only
    int32_t* fe = nullptr;
    uint32_t* ve_off = nullptr;
    int32_t* ve = nullptr;
    uint32_t* ef_off = nullptr;
    int32_t* ef = nullptr;
    int* err = nullptr;
    uint32_t* sums = nullptr;
    bool degenerate = false; // an edge with equal ends or a repeated side
    Mem mem;
};

// the state of CellComplexThinning over one vc_cc: measures, counts, the
// generation of every removal and the queue buffers (grown on demand and
// kept across prune calls)
struct vc_ccthin
{
    vc_cc* cc = nullptr;
    float* em = nullptr;
    float* fm = nullptr;
    int32_t *ref_e = nullptr, *ref_v = nullptr;
    int32_t *rg_f = nullptr, *rj_f = nullptr, *rg_e = nullptr, *rj_e = nullptr;
    int32_t *first_f = nullptr, *first_e = nullptr;
    uint8_t *rm_v = nullptr, *rm_e = nullptr, *rm_f = nullptr;
    uint32_t* qflag = nullptr;
    bool pruned = false;
    size_t cap_p = 0, cap_s = 0;
    Pairs P{}, N{}, SL{};
    uint8_t* eff = nullptr;
    uint32_t *slot_off = nullptr, *sflag = nullptr;
    Mem mem;
    Mem grown;
};

namespace
{

// adjacency and validation of a complex whose E, fo, fe are on the device
int cc_adjacency(vc_ctx* ctx, vc_cc* c)
{
    hipStream_t S = ctx->stream;
    const size_t V = (size_t)c->nv + 1, E = (size_t)c->ne + 1;
    int rc = VC_OK;
    uint32_t *fillv = nullptr, *fille = nullptr;
    Mem tmp;
    auto get = [&](Mem& m, auto** p, size_t n) {
        if (rc == VC_OK)
            rc = m.get(ctx, p, n);
    };
    get(c->mem, &c->ve_off, V + 1);
    get(c->mem, &c->ef_off, E + 1);
    get(c->mem, &c->ve, 2 * E);
    get(c->mem, &c->ef, (size_t)c->nfe + 1);
    get(c->mem, &c->err, 4);
    get(c->mem, &c->sums, (size_t)kScanB + 1);
    get(tmp, &fillv, V);
    get(tmp, &fille, E);
    if (rc != VC_OK)
        return rc;
    VC_HIP(ctx, hipMemsetAsync(c->ve_off, 0, 4 * (V + 1), S));
    VC_HIP(ctx, hipMemsetAsync(c->ef_off, 0, 4 * (E + 1), S));
    VC_HIP(ctx, hipMemsetAsync(fillv, 0, 4 * V, S));
    VC_HIP(ctx, hipMemsetAsync(fille, 0, 4 * E, S));
    VC_HIP(ctx, hipMemsetAsync(c->err, 0, sizeof(int), S));
    if (c->ne)
        deg_v<<<nblk(c->ne), kB, 0, S>>>(c->E, c->ne, c->nv, c->ve_off, c->err);
    if (c->nf)
        deg_e<<<nblk(c->nf), kB, 0, S>>>(c->fo, c->fe, c->nf, c->ne, c->ef_off, c->err);
    VC_CHECK_LAUNCH(ctx);
    int herr = 0;
    if ((rc = read_err(ctx, c->err, &herr)) != VC_OK)
        return rc;
    if (herr & kAdjRange)
        return vc_fail(ctx, VC_ERR_INVALID,
                       "cell complex: an edge end or a face side out of range");
    c->degenerate = (herr & kAdjDegenerate) != 0;
    if ((rc = scan_any(ctx, c->ve_off, (int64_t)V, c->sums)) != VC_OK ||
        (rc = scan_any(ctx, c->ef_off, (int64_t)E, c->sums)) != VC_OK)
        return rc;
    if (c->ne)
        fill_ve<<<nblk(c->ne), kB, 0, S>>>(c->E, c->ne, c->ve_off, fillv, c->ve);
    if (c->nf)
        fill_ef<<<nblk(c->nf), kB, 0, S>>>(c->fo, c->fe, c->nf, c->ef_off, fille, c->ef);
    VC_CHECK_LAUNCH(ctx);
    if (c->nv)
        sort_lists<<<nblk(c->nv), kB, 0, S>>>(c->ve_off, c->nv, c->ve);
    if (c->ne)
        sort_lists<<<nblk(c->ne), kB, 0, S>>>(c->ef_off, c->ne, c->ef);
    VC_CHECK_LAUNCH(ctx);
    // fillv / fille are freed on return
    VC_HIP(ctx, hipStreamSynchronize(S));
    return VC_OK;
}

int thin_alloc(vc_ctx* ctx, vc_ccthin* t)
{
    const vc_cc* c = t->cc;
    const size_t V = (size_t)c->nv + 1, E = (size_t)c->ne + 1, F = (size_t)c->nf + 1;
    int rc = VC_OK;
    auto get = [&](auto** p, size_t n) {
        if (rc == VC_OK)
            rc = t->mem.get(ctx, p, n);
    };
    get(&t->em, E);
    get(&t->fm, F);
    get(&t->ref_e, E);
    get(&t->ref_v, V);
    get(&t->rg_f, F);
    get(&t->rj_f, F);
    get(&t->rg_e, E);
    get(&t->rj_e, E);
    get(&t->first_f, F);
    get(&t->first_e, E);
    get(&t->rm_v, V);
    get(&t->rm_e, E);
    get(&t->rm_f, F);
    get(&t->qflag, (size_t)c->ne + c->nv + 1);
    if (rc != VC_OK)
        return rc;
    hipStream_t S = ctx->stream;
    VC_HIP(ctx, hipMemsetAsync(t->em, 0, 4 * E, S));
    VC_HIP(ctx, hipMemsetAsync(t->fm, 0, 4 * F, S));
    return VC_OK;
}

// the pair buffers hold at least n pairs (the current ones are kept)
int grow_pairs(vc_ctx* ctx, vc_ccthin* t, size_t n)
{
    if (n <= t->cap_p)
        return VC_OK;
    hipStream_t S = ctx->stream;
    const size_t c = n + n / 2 + 1024;
    uint8_t *t1, *t2, *e1;
    int32_t *c1, *o1, *c2, *o2;
    uint32_t* s1;
    int rc = VC_OK;
    auto get = [&](auto** p, size_t k) {
        if (rc == VC_OK)
            rc = t->grown.get(ctx, p, k);
    };
    get(&t1, c);
    get(&t2, c);
    get(&e1, c);
    get(&c1, c);
    get(&o1, c);
    get(&c2, c);
    get(&o2, c);
    get(&s1, c);
    if (rc != VC_OK)
        return rc;
    if (t->cap_p)
    {
        VC_HIP(ctx, hipMemcpyAsync(t1, t->P.t, t->cap_p, hipMemcpyDeviceToDevice, S));
        VC_HIP(ctx, hipMemcpyAsync(c1, t->P.c, 4 * t->cap_p, hipMemcpyDeviceToDevice, S));
        VC_HIP(ctx, hipMemcpyAsync(o1, t->P.o, 4 * t->cap_p, hipMemcpyDeviceToDevice, S));
    }
    t->P = Pairs{t1, c1, o1};
    t->N = Pairs{t2, c2, o2};
    t->eff = e1;
    t->slot_off = s1;
    t->cap_p = c;
    return VC_OK;
}

int grow_slots(vc_ctx* ctx, vc_ccthin* t, size_t n)
{
    if (n <= t->cap_s)
        return VC_OK;
    const size_t c = n + n / 2 + 1024;
    uint8_t* tt;
    int32_t *cc, *oo;
    uint32_t* fl;
    int rc = VC_OK;
    auto get = [&](auto** p, size_t k) {
        if (rc == VC_OK)
            rc = t->grown.get(ctx, p, k);
    };
    get(&tt, c);
    get(&cc, c);
    get(&oo, c);
    get(&fl, c);
    if (rc != VC_OK)
        return rc;
    t->SL = Pairs{tt, cc, oo};
    t->sflag = fl;
    t->cap_s = c;
    return VC_OK;
}

// CellComplexThinning::prune(f_t, l_t) on the resident complex: every call
// starts from the full complex (src/ccthin.cpp:201-424); the results stay
// on the device (rm_v, rm_e, rm_f, ref_e)
int thin_prune(vc_ctx* ctx, vc_ccthin* t, float f_t, float l_t, int64_t* gens)
{
    const vc_cc* c = t->cc;
    const int32_t nv = c->nv, ne = c->ne, nf = c->nf;
    hipStream_t S = ctx->stream;
    if (gens)
        *gens = 0;
    t->pruned = false;
    if (c->degenerate)
        return vc_fail(ctx, VC_ERR_INVALID,
                       "collapse on the device: the complex has an edge with "
                       "equal ends or a face listing a side twice");
    const int32_t mx = std::max(nv, std::max(ne, nf));
    VC_HIP(ctx, hipMemsetAsync(c->err, 0, sizeof(int), S));
    if (mx)
        reset_state<<<nblk(mx), kB, 0, S>>>(nv, ne, nf, c->ve_off, c->ef_off,
                                             t->ref_v, t->ref_e, t->rg_f, t->rg_e,
                                             t->first_f, t->first_e, t->rm_v);
    VC_CHECK_LAUNCH(ctx);
    Cx C{};
    C.E = c->E;
    C.fo = c->fo;
    C.fe = c->fe;
    C.ve_off = c->ve_off;
    C.ve = c->ve;
    C.ef_off = c->ef_off;
    C.ef = c->ef;
    C.em = t->em;
    C.fm = t->fm;
    C.ref_e = t->ref_e;
    C.ref_v = t->ref_v;
    C.rg_f = t->rg_f;
    C.rj_f = t->rj_f;
    C.rg_e = t->rg_e;
    C.rj_e = t->rj_e;
    C.rm_v = t->rm_v;
    C.first_f = t->first_f;
    C.first_e = t->first_e;
    C.err = c->err;

    // generation 0: face-edge pairs by edge, then edge-vertex pairs by vertex
    // (src/ccthin.cpp:248-270)
    const int64_t nq = (int64_t)ne + nv;
    uint32_t* qflag = t->qflag;
    VC_HIP(ctx, hipMemsetAsync(qflag, 0, 4 * ((size_t)nq + 1), S));
    if (nq)
        init_flags<<<nblk(nq), kB, 0, S>>>(ne, nv, t->ref_e, t->ref_v, c->ef_off,
                                            c->ef, c->ve_off, c->ve, t->em, t->fm,
                                            f_t, l_t, qflag);
    VC_CHECK_LAUNCH(ctx);
    int rc = scan_any(ctx, qflag, nq + 1, c->sums);
    if (rc != VC_OK)
        return rc;
    uint32_t np = 0;
    VC_HIP(ctx, hipMemcpyAsync(&np, qflag + nq, 4, hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    if ((rc = grow_pairs(ctx, t, (size_t)np + 1)) != VC_OK)
        return rc;
    if (np)
        init_pairs<<<nblk(nq), kB, 0, S>>>(ne, nv, qflag, qflag + nq, c->ef_off,
                                            c->ef, c->ve_off, c->ve, t->P);
    VC_CHECK_LAUNCH(ctx);

    int32_t g = 0;
    while (np > 0 && rc == VC_OK)
    {
        const int32_t n = (int32_t)np;
        first_face<<<nblk(n), kB, 0, S>>>(C, t->P, n);
        first_edge<<<nblk(n), kB, 0, S>>>(C, t->P, n, t->eff, t->slot_off);
        decide<<<nblk(n), kB, 0, S>>>(C, t->P, n, t->eff);
        VC_CHECK_LAUNCH(ctx);
        VC_HIP(ctx, hipMemsetAsync(t->slot_off + n, 0, 4, S));
        if ((rc = scan_any(ctx, t->slot_off, (int64_t)n + 1, c->sums)) != VC_OK)
            break;
        uint32_t hs[2] = {0, 0};
        VC_HIP(ctx, hipMemcpyAsync(&hs[0], t->slot_off + n, 4, hipMemcpyDeviceToHost, S));
        VC_HIP(ctx, hipMemcpyAsync(&hs[1], c->err, 4, hipMemcpyDeviceToHost, S));
        VC_HIP(ctx, hipStreamSynchronize(S));
        if (hs[1])
        {
            rc = vc_fail(ctx, VC_ERR_STATE,
                         hs[1] == kErrLogic
                             ? "collapse: a popped pair's count is not 1 (logic err)"
                             : "collapse: an edge removed twice in one generation");
            break;
        }
        const uint32_t ns = hs[0];
        if ((rc = grow_slots(ctx, t, (size_t)ns + 1)) != VC_OK)
            break;
        apply<<<nblk(n), kB, 0, S>>>(C, t->P, n, t->eff, g);
        fill_u32<<<nblk((int64_t)ns + 1), kB, 0, S>>>(t->sflag, (int64_t)ns + 1, 0u);
        trigger<<<nblk(n), kB, 0, S>>>(C, t->P, n, t->eff, g, t->slot_off, f_t, l_t,
                                         t->sflag, t->SL);
        VC_CHECK_LAUNCH(ctx);
        if ((rc = scan_any(ctx, t->sflag, (int64_t)ns + 1, c->sums)) != VC_OK)
            break;
        uint32_t nn = 0;
        VC_HIP(ctx, hipMemcpyAsync(&nn, t->sflag + ns, 4, hipMemcpyDeviceToHost, S));
        VC_HIP(ctx, hipStreamSynchronize(S));
        if ((rc = grow_pairs(ctx, t, (size_t)nn + 1)) != VC_OK)
            break;
        if (nn)
            compact<<<nblk(ns), kB, 0, S>>>(t->sflag, ns, t->sflag + ns, t->SL, t->N);
        VC_CHECK_LAUNCH(ctx);
        std::swap(t->P, t->N);
        np = nn;
        ++g;
    }
    if (rc != VC_OK)
        return rc;
    const int32_t me = ne > nf ? ne : nf;
    if (me)
        results<<<nblk(me), kB, 0, S>>>(ne, nf, t->rg_e, t->rg_f, t->rm_e, t->rm_f);
    VC_CHECK_LAUNCH(ctx);
    t->pruned = true;
    if (gens)
        *gens = g;
    return VC_OK;
}

// the remaining complex as flags scanned into positions; counts in *nv_out,
// *ne_out, *nt_out; the position arrays are allocated in m
int thin_remaining_pos(vc_ctx* ctx, vc_ccthin* t, Mem& m, uint32_t** vpos,
                       uint32_t** epos, uint32_t** tpos, int64_t* nv_out,
                       int64_t* ne_out, int64_t* nt_out)
{
    const vc_cc* c = t->cc;
    hipStream_t S = ctx->stream;
    int rc = VC_OK;
    auto get = [&](auto** p, size_t n) {
        if (rc == VC_OK)
            rc = m.get(ctx, p, n);
    };
    get(vpos, (size_t)c->nv + 1);
    get(epos, (size_t)c->ne + 1);
    get(tpos, (size_t)c->nf + 1);
    if (rc != VC_OK)
        return rc;
    VC_HIP(ctx, hipMemsetAsync(*vpos + c->nv, 0, 4, S));
    VC_HIP(ctx, hipMemsetAsync(*epos + c->ne, 0, 4, S));
    VC_HIP(ctx, hipMemsetAsync(*tpos + c->nf, 0, 4, S));
    const int32_t mx = std::max(c->nv, std::max(c->ne, c->nf));
    if (mx)
        rem_flags<<<nblk(mx), kB, 0, S>>>(c->nv, c->ne, c->nf, t->rm_v, t->rm_e,
                                           t->ref_e, t->rm_f, c->fo, *vpos, *epos,
                                           *tpos);
    VC_CHECK_LAUNCH(ctx);
    if ((rc = scan_any(ctx, *vpos, (int64_t)c->nv + 1, c->sums)) != VC_OK ||
        (rc = scan_any(ctx, *epos, (int64_t)c->ne + 1, c->sums)) != VC_OK ||
        (rc = scan_any(ctx, *tpos, (int64_t)c->nf + 1, c->sums)) != VC_OK)
        return rc;
    uint32_t h[3];
    VC_HIP(ctx, hipMemcpyAsync(&h[0], *vpos + c->nv, 4, hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipMemcpyAsync(&h[1], *epos + c->ne, 4, hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipMemcpyAsync(&h[2], *tpos + c->nf, 4, hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    *nv_out = h[0];
    *ne_out = h[1];
    *nt_out = h[2];
    return VC_OK;
}

// the complex's face components (compute_conn_cmpnts) or vertex components
// through its edges (findNumConnComponents), by union-find
int cc_components(vc_ctx* ctx, vc_cc* c, bool faces, int64_t* out)
{
    hipStream_t S = ctx->stream;
    const int64_t n = faces ? c->nf : c->nv;
    Mem m;
    int32_t* par = nullptr;
    unsigned long long* cnt = nullptr;
    int rc = m.get(ctx, &par, (size_t)n + 1);
    if (rc == VC_OK)
        rc = m.get(ctx, &cnt, 1);
    if (rc != VC_OK)
        return rc;
    VC_HIP(ctx, hipMemsetAsync(c->err, 0, sizeof(int), S));
    if (n)
        uf_init<<<nblk(n), kB, 0, S>>>(par, n);
    if (c->ne)
    {
        if (faces)
            uf_faces<<<nblk(c->ne), kB, 0, S>>>(c->ef_off, c->ef, c->ne, par, c->err);
        else
            uf_pairs<<<nblk(c->ne), kB, 0, S>>>(c->E, c->ne, par, c->err);
    }
    VC_CHECK_LAUNCH(ctx);
    int herr = 0;
    if ((rc = count_roots(ctx, par, n, cnt, out)) != VC_OK ||
        (rc = read_err(ctx, c->err, &herr)) != VC_OK)
        return rc;
    if (herr)
        return vc_fail(ctx, VC_ERR_GUARD, "union-find: iteration guard tripped");
    return VC_OK;
}

// util::computeEulerChar(vts, edges, tris) on device arrays (src/geomalgo.cpp:
// 174-214): V = nv, E = distinct pairs among the edges as given and the
// normalised triangle sides, F = nt, C = vertex components through them
int euler_arrays_dev(vc_ctx* ctx, int32_t nv, const int32_t* dE, int64_t ne,
                     const int32_t* dT, int64_t nt, int* derr, vc_cc_topo* out)
{
    hipStream_t S = ctx->stream;
    const int64_t nc = ne + 3 * nt;
    const uint64_t H = table_size(nc);
    Mem m;
    int32_t *cand = nullptr, *par = nullptr;
    uint64_t* keys = nullptr;
    unsigned long long *vals = nullptr, *cnt = nullptr;
    int rc = VC_OK;
    auto get = [&](auto** p, size_t n) {
        if (rc == VC_OK)
            rc = m.get(ctx, p, n);
    };
    get(&cand, 2 * (size_t)nc + 2);
    get(&par, (size_t)nv + 1);
    get(&keys, (size_t)H);
    get(&vals, (size_t)H);
    get(&cnt, 1);
    if (rc != VC_OK)
        return rc;
    fill_u64<<<nblk((int64_t)H), kB, 0, S>>>(keys, (int64_t)H, kEmpty);
    fill_u64<<<nblk((int64_t)H), kB, 0, S>>>(reinterpret_cast<uint64_t*>(vals),
                                              (int64_t)H, ~0ull);
    VC_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(unsigned long long), S));
    VC_HIP(ctx, hipMemsetAsync(derr, 0, sizeof(int), S));
    if (nc)
        euler_cands<<<nblk(nc), kB, 0, S>>>(dE, ne, dT, nt, cand, keys, vals, H - 1);
    if (nc)
        count_first<<<nblk(nc), kB, 0, S>>>(cand, nc, keys, vals, H - 1, cnt);
    if (nv)
        uf_init<<<nblk(nv), kB, 0, S>>>(par, nv);
    if (nc)
        uf_pairs<<<nblk(nc), kB, 0, S>>>(cand, nc, par, derr);
    VC_CHECK_LAUNCH(ctx);
    unsigned long long nE = 0;
    VC_HIP(ctx, hipMemcpyAsync(&nE, cnt, sizeof(nE), hipMemcpyDeviceToHost, S));
    int64_t C = 0;
    int herr = 0;
    if ((rc = count_roots(ctx, par, nv, cnt, &C)) != VC_OK ||
        (rc = read_err(ctx, derr, &herr)) != VC_OK)
        return rc;
    if (herr)
        return vc_fail(ctx, VC_ERR_GUARD, "union-find: iteration guard tripped");
    std::memset(out, 0, sizeof(*out));
    out->V = nv;
    out->E = (int64_t)nE;
    out->F = nt;
    out->T = 0;
    out->C = C;
    out->euler = out->V - out->E + out->F;
    out->face_components = -1;
    return VC_OK;
}

int check_handle(vc_ctx* ctx, const vc_cc* c)
{
    return ctx && c && c->ctx == ctx ? VC_OK : VC_ERR_INVALID;
}

} // namespace

extern "C" {

int vc_cc_create(vc_ctx* ctx, int32_t nv, int32_t ne_in, const int32_t* edges_in,
                 int32_t nf, const int64_t* face_off, const int32_t* face_vts,
                 vc_cc** out)
{
    if (!ctx || !out || nv < 0 || ne_in < 0 || nf < 0 || (ne_in && !edges_in) ||
        (nf && (!face_off || !face_vts)))
        return VC_ERR_INVALID;
    *out = nullptr;
    int rc = vc_cc_check_host(ctx, nv, ne_in, edges_in, nf, face_off, face_vts);
    if (rc != VC_OK)
        return rc;
    const int64_t nfv = nf ? face_off[nf] : 0;
    if (nfv >= (int64_t)UINT32_MAX / 2 || 2 * ((int64_t)ne_in + nfv) >= (int64_t)INT32_MAX)
        return vc_fail(ctx, VC_ERR_INVALID, "vc_cc_create: complex too large");
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    vc_cc* c = new (std::nothrow) vc_cc();
    if (!c)
        return VC_ERR_NOMEM;
    c->ctx = ctx;
    c->nv = nv;
    c->nf = nf;
    c->nfe = nfv;
    int32_t* dEin = nullptr;
    Mem tmp;
    auto get = [&](Mem& m, auto** p, size_t n) {
        if (rc == VC_OK)
            rc = m.get(ctx, p, n);
    };
    get(tmp, &dEin, 2 * (size_t)ne_in + 2);
    get(c->mem, &c->fo, (size_t)nf + 1);
    get(c->mem, &c->fv, (size_t)nfv + 1);
    get(c->mem, &c->fe, (size_t)nfv + 1);
    get(c->mem, &c->E, 2 * ((size_t)ne_in + nfv) + 2);
    auto fail = [&](int code) {
        delete c;
        return code;
    };
    if (rc != VC_OK)
        return fail(rc);
    hipError_t e = hipSuccess;
    if (ne_in)
        e = hipMemcpyAsync(dEin, edges_in, 8 * (size_t)ne_in, hipMemcpyHostToDevice, S);
    if (e == hipSuccess && nf)
        e = hipMemcpyAsync(c->fo, face_off, 8 * ((size_t)nf + 1), hipMemcpyHostToDevice, S);
    if (e == hipSuccess && nfv)
        e = hipMemcpyAsync(c->fv, face_vts, 4 * (size_t)nfv, hipMemcpyHostToDevice, S);
    if (e != hipSuccess)
        return fail(vc_fail(ctx, VC_ERR_HIP, "vc_cc_create: %s", hipGetErrorString(e)));
    int32_t ne = 0;
    if ((rc = vc_cc_finalize_dev(ctx, ne_in, dEin, nf, c->fo, c->fv, nfv, c->E, c->fe,
                                 &ne)) != VC_OK)
        return fail(rc);
    c->ne = ne;
    if ((rc = cc_adjacency(ctx, c)) != VC_OK)
        return fail(rc);
    *out = c;
    return VC_OK;
}

int vc_cc_release(vc_ctx* ctx, vc_cc* c)
{
    if (check_handle(ctx, c) != VC_OK)
        return VC_ERR_INVALID;
    (void)hipStreamSynchronize(ctx->stream);
    delete c;
    return VC_OK;
}

int vc_cc_sizes(const vc_cc* c, int32_t* nv, int32_t* ne, int32_t* nf,
                int64_t* n_sides)
{
    if (!c)
        return VC_ERR_INVALID;
    if (nv)
        *nv = c->nv;
    if (ne)
        *ne = c->ne;
    if (nf)
        *nf = c->nf;
    if (n_sides)
        *n_sides = c->nfe;
    return VC_OK;
}

int vc_cc_get(vc_ctx* ctx, vc_cc* c, int32_t* edges, int32_t* face_edges)
{
    if (check_handle(ctx, c) != VC_OK)
        return VC_ERR_INVALID;
    hipStream_t S = ctx->stream;
    if (edges && c->ne)
        VC_HIP(ctx, hipMemcpyAsync(edges, c->E, 8 * (size_t)c->ne,
                                   hipMemcpyDeviceToHost, S));
    if (face_edges && c->nfe)
        VC_HIP(ctx, hipMemcpyAsync(face_edges, c->fe, 4 * (size_t)c->nfe,
                                   hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    return VC_OK;
}

int vc_cc_topology(vc_ctx* ctx, vc_cc* c, vc_cc_topo* out)
{
    if (check_handle(ctx, c) != VC_OK || !out)
        return VC_ERR_INVALID;
    VC_HIP(ctx, hipSetDevice(ctx->device));
    int64_t C = 0, FC = 0;
    int rc = cc_components(ctx, c, false, &C);
    if (rc == VC_OK)
        rc = cc_components(ctx, c, true, &FC);
    if (rc != VC_OK)
        return rc;
    std::memset(out, 0, sizeof(*out));
    out->V = c->nv;
    out->E = c->ne;
    out->F = c->nf;
    out->T = 0;
    out->euler = out->V - out->E + out->F;
    out->C = C;
    out->face_components = FC;
    return VC_OK;
}

int vc_cc_subset_components(vc_ctx* ctx, vc_cc* c, const uint8_t* face_in,
                             int64_t* n_comp)
{
    if (check_handle(ctx, c) != VC_OK || !n_comp || (c->nf && !face_in))
        return VC_ERR_INVALID;
    if (c->nf && !c->fv)
        return vc_fail(ctx, VC_ERR_STATE, "vc_cc_subset_components: no face vertices");
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    Mem m;
    uint8_t *din = nullptr, *in_set = nullptr;
    int32_t* par = nullptr;
    unsigned long long* cnt = nullptr;
    int rc = VC_OK;
    auto get = [&](auto** p, size_t n) {
        if (rc == VC_OK)
            rc = m.get(ctx, p, n);
    };
    get(&din, (size_t)c->nf + 1);
    get(&in_set, (size_t)c->nv + 1);
    get(&par, (size_t)c->nv + 1);
    get(&cnt, 1);
    if (rc != VC_OK)
        return rc;
    if (c->nf)
        VC_HIP(ctx, hipMemcpyAsync(din, face_in, c->nf, hipMemcpyHostToDevice, S));
    VC_HIP(ctx, hipMemsetAsync(in_set, 0, (size_t)c->nv + 1, S));
    VC_HIP(ctx, hipMemsetAsync(c->err, 0, sizeof(int), S));
    VC_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(unsigned long long), S));
    if (c->nv)
        uf_init<<<nblk(c->nv), kB, 0, S>>>(par, c->nv);
    if (c->nf)
        subset_union<<<nblk(c->nf), kB, 0, S>>>(c->nf, c->fo, c->fv, din, in_set, par,
                                                 c->err);
    if (c->nv)
        uf_roots_in<<<nblk(c->nv), kB, 0, S>>>(par, in_set, c->nv, cnt);
    VC_CHECK_LAUNCH(ctx);
    unsigned long long h = 0;
    int herr = 0;
    VC_HIP(ctx, hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, S));
    if ((rc = read_err(ctx, c->err, &herr)) != VC_OK)
        return rc;
    if (herr)
        return vc_fail(ctx, VC_ERR_GUARD, "union-find: iteration guard tripped");
    *n_comp = (int64_t)h;
    return VC_OK;
}

int vc_ccthin_create(vc_ctx* ctx, vc_cc* c, vc_ccthin** out)
{
    if (check_handle(ctx, c) != VC_OK || !out)
        return VC_ERR_INVALID;
    *out = nullptr;
    VC_HIP(ctx, hipSetDevice(ctx->device));
    vc_ccthin* t = new (std::nothrow) vc_ccthin();
    if (!t)
        return VC_ERR_NOMEM;
    t->cc = c;
    const int rc = thin_alloc(ctx, t);
    if (rc != VC_OK)
    {
        delete t;
        return rc;
    }
    *out = t;
    return VC_OK;
}

int vc_ccthin_release(vc_ctx* ctx, vc_ccthin* t)
{
    if (!t || check_handle(ctx, t->cc) != VC_OK)
        return VC_ERR_INVALID;
    (void)hipStreamSynchronize(ctx->stream);
    delete t;
    return VC_OK;
}

int vc_ccthin_assign(vc_ctx* ctx, vc_ccthin* t, const float* edge_measure,
                     const float* face_measure)
{
    if (!t || check_handle(ctx, t->cc) != VC_OK ||
        (t->cc->ne && !edge_measure) || (t->cc->nf && !face_measure))
        return VC_ERR_INVALID;
    hipStream_t S = ctx->stream;
    if (t->cc->ne)
        VC_HIP(ctx, hipMemcpyAsync(t->em, edge_measure, 4 * (size_t)t->cc->ne,
                                   hipMemcpyHostToDevice, S));
    if (t->cc->nf)
        VC_HIP(ctx, hipMemcpyAsync(t->fm, face_measure, 4 * (size_t)t->cc->nf,
                                   hipMemcpyHostToDevice, S));
    // the caller's arrays may be released on return
    VC_HIP(ctx, hipStreamSynchronize(S));
    return VC_OK;
}

int vc_ccthin_prune(vc_ctx* ctx, vc_ccthin* t, float f_t, float l_t,
                    int64_t* generations)
{
    if (!t || check_handle(ctx, t->cc) != VC_OK)
        return VC_ERR_INVALID;
    VC_HIP(ctx, hipSetDevice(ctx->device));
    return thin_prune(ctx, t, f_t, l_t, generations);
}

int vc_ccthin_removed(vc_ctx* ctx, vc_ccthin* t, uint8_t* rv, uint8_t* re,
                      uint8_t* rf, int32_t* ref_e)
{
    if (!t || check_handle(ctx, t->cc) != VC_OK)
        return VC_ERR_INVALID;
    if (!t->pruned)
        return vc_fail(ctx, VC_ERR_STATE, "vc_ccthin_removed: prune first");
    const vc_cc* c = t->cc;
    hipStream_t S = ctx->stream;
    if (rv && c->nv)
        VC_HIP(ctx, hipMemcpyAsync(rv, t->rm_v, c->nv, hipMemcpyDeviceToHost, S));
    if (re && c->ne)
        VC_HIP(ctx, hipMemcpyAsync(re, t->rm_e, c->ne, hipMemcpyDeviceToHost, S));
    if (rf && c->nf)
        VC_HIP(ctx, hipMemcpyAsync(rf, t->rm_f, c->nf, hipMemcpyDeviceToHost, S));
    if (ref_e && c->ne)
        VC_HIP(ctx, hipMemcpyAsync(ref_e, t->ref_e, 4 * (size_t)c->ne,
                                   hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    return VC_OK;
}

int vc_ccthin_remaining(vc_ctx* ctx, vc_ccthin* t, int32_t* vts_idx,
                        int64_t vts_cap, int64_t* n_vts, int32_t* edges,
                        int64_t edge_cap, int64_t* n_edges, int32_t* tris,
                        int32_t* from_fi, int64_t tri_cap, int64_t* n_tris)
{
    if (!t || check_handle(ctx, t->cc) != VC_OK || !n_vts || !n_edges || !n_tris)
        return VC_ERR_INVALID;
    if (!t->pruned)
        return vc_fail(ctx, VC_ERR_STATE, "vc_ccthin_remaining: prune first");
    const vc_cc* c = t->cc;
    if (c->nf && !c->fv)
        return vc_fail(ctx, VC_ERR_STATE, "vc_ccthin_remaining: no face vertices");
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    Mem m;
    uint32_t *vpos, *epos, *tpos;
    int64_t NV = 0, NE = 0, NT = 0;
    int rc = thin_remaining_pos(ctx, t, m, &vpos, &epos, &tpos, &NV, &NE, &NT);
    if (rc != VC_OK)
        return rc;
    *n_vts = NV;
    *n_edges = NE;
    *n_tris = NT;
    const bool query = !vts_idx && !edges && !tris && !from_fi;
    if (query)
        return VC_OK;
    if ((vts_idx && NV > vts_cap) || (edges && NE > edge_cap) ||
        ((tris || from_fi) && NT > tri_cap))
        return VC_ERR_CAPACITY;
    int32_t *dv = nullptr, *de = nullptr, *dt = nullptr, *df = nullptr;
    auto get = [&](auto** p, size_t n) {
        if (rc == VC_OK)
            rc = m.get(ctx, p, n);
    };
    if (vts_idx)
        get(&dv, (size_t)NV + 1);
    if (edges)
        get(&de, 2 * (size_t)NE + 2);
    if (tris)
        get(&dt, 3 * (size_t)NT + 3);
    if (from_fi)
        get(&df, (size_t)NT + 1);
    if (rc != VC_OK)
        return rc;
    const int32_t mx = std::max(c->nv, std::max(c->ne, c->nf));
    if (mx)
        rem_emit<<<nblk(mx), kB, 0, S>>>(c->nv, c->ne, c->nf, c->E, c->fo, c->fv, vpos,
                                          epos, tpos, dv, de, dt, df);
    VC_CHECK_LAUNCH(ctx);
    if (dv && NV)
        VC_HIP(ctx, hipMemcpyAsync(vts_idx, dv, 4 * (size_t)NV, hipMemcpyDeviceToHost, S));
    if (de && NE)
        VC_HIP(ctx, hipMemcpyAsync(edges, de, 8 * (size_t)NE, hipMemcpyDeviceToHost, S));
    if (dt && NT)
        VC_HIP(ctx, hipMemcpyAsync(tris, dt, 12 * (size_t)NT, hipMemcpyDeviceToHost, S));
    if (df && NT)
        VC_HIP(ctx, hipMemcpyAsync(from_fi, df, 4 * (size_t)NT, hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    return VC_OK;
}

int vc_ccthin_remained(vc_ctx* ctx, vc_ccthin* t, int32_t* v_ids, int64_t v_cap,
                       int64_t* n_v, int32_t* e_ids, int64_t e_cap, int64_t* n_e,
                       int32_t* f_ids, int64_t f_cap, int64_t* n_f)
{
    if (!t || check_handle(ctx, t->cc) != VC_OK || !n_v || !n_e || !n_f)
        return VC_ERR_INVALID;
    if (!t->pruned)
        return vc_fail(ctx, VC_ERR_STATE, "vc_ccthin_remained: prune first");
    const vc_cc* c = t->cc;
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    Mem m;
    uint32_t *vpos = nullptr, *epos = nullptr, *fpos = nullptr;
    int rc = VC_OK;
    auto get = [&](auto** p, size_t n) {
        if (rc == VC_OK)
            rc = m.get(ctx, p, n);
    };
    get(&vpos, (size_t)c->nv + 1);
    get(&epos, (size_t)c->ne + 1);
    get(&fpos, (size_t)c->nf + 1);
    if (rc != VC_OK)
        return rc;
    VC_HIP(ctx, hipMemsetAsync(vpos + c->nv, 0, 4, S));
    VC_HIP(ctx, hipMemsetAsync(epos + c->ne, 0, 4, S));
    VC_HIP(ctx, hipMemsetAsync(fpos + c->nf, 0, 4, S));
    const int32_t mx = std::max(c->nv, std::max(c->ne, c->nf));
    if (mx)
        kept_flags<<<nblk(mx), kB, 0, S>>>(c->nv, c->ne, c->nf, t->rm_v, t->rm_e,
                                            t->ref_e, t->rm_f, vpos, epos, fpos);
    VC_CHECK_LAUNCH(ctx);
    if ((rc = scan_any(ctx, vpos, (int64_t)c->nv + 1, c->sums)) != VC_OK ||
        (rc = scan_any(ctx, epos, (int64_t)c->ne + 1, c->sums)) != VC_OK ||
        (rc = scan_any(ctx, fpos, (int64_t)c->nf + 1, c->sums)) != VC_OK)
        return rc;
    uint32_t h[3];
    VC_HIP(ctx, hipMemcpyAsync(&h[0], vpos + c->nv, 4, hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipMemcpyAsync(&h[1], epos + c->ne, 4, hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipMemcpyAsync(&h[2], fpos + c->nf, 4, hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    *n_v = h[0];
    *n_e = h[1];
    *n_f = h[2];
    if (!v_ids && !e_ids && !f_ids)
        return VC_OK;
    if ((v_ids && *n_v > v_cap) || (e_ids && *n_e > e_cap) || (f_ids && *n_f > f_cap))
        return VC_ERR_CAPACITY;
    int32_t *dv = nullptr, *de = nullptr, *df = nullptr;
    if (v_ids)
        get(&dv, (size_t)*n_v + 1);
    if (e_ids)
        get(&de, (size_t)*n_e + 1);
    if (f_ids)
        get(&df, (size_t)*n_f + 1);
    if (rc != VC_OK)
        return rc;
    if (mx)
        kept_ids<<<nblk(mx), kB, 0, S>>>(c->nv, c->ne, c->nf, vpos, epos, fpos, dv, de,
                                          df);
    VC_CHECK_LAUNCH(ctx);
    if (dv && *n_v)
        VC_HIP(ctx, hipMemcpyAsync(v_ids, dv, 4 * (size_t)*n_v, hipMemcpyDeviceToHost, S));
    if (de && *n_e)
        VC_HIP(ctx, hipMemcpyAsync(e_ids, de, 4 * (size_t)*n_e, hipMemcpyDeviceToHost, S));
    if (df && *n_f)
        VC_HIP(ctx, hipMemcpyAsync(f_ids, df, 4 * (size_t)*n_f, hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    return VC_OK;
}

int vc_ccthin_remaining_euler(vc_ctx* ctx, vc_ccthin* t, vc_cc_topo* out)
{
    if (!t || check_handle(ctx, t->cc) != VC_OK || !out)
        return VC_ERR_INVALID;
    if (!t->pruned)
        return vc_fail(ctx, VC_ERR_STATE, "vc_ccthin_remaining_euler: prune first");
    const vc_cc* c = t->cc;
    if (c->nf && !c->fv)
        return vc_fail(ctx, VC_ERR_STATE, "vc_ccthin_remaining_euler: no face vertices");
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    Mem m;
    uint32_t *vpos, *epos, *tpos;
    int64_t NV = 0, NE = 0, NT = 0;
    int rc = thin_remaining_pos(ctx, t, m, &vpos, &epos, &tpos, &NV, &NE, &NT);
    if (rc != VC_OK)
        return rc;
    int32_t *de = nullptr, *dt = nullptr;
    if ((rc = m.get(ctx, &de, 2 * (size_t)NE + 2)) != VC_OK ||
        (rc = m.get(ctx, &dt, 3 * (size_t)NT + 3)) != VC_OK)
        return rc;
    const int32_t mx = std::max(c->nv, std::max(c->ne, c->nf));
    if (mx)
        rem_emit<<<nblk(mx), kB, 0, S>>>(c->nv, c->ne, c->nf, c->E, c->fo, c->fv, vpos,
                                          epos, tpos, nullptr, de, dt, nullptr);
    VC_CHECK_LAUNCH(ctx);
    return euler_arrays_dev(ctx, (int32_t)NV, de, NE, dt, NT, c->err, out);
}

int vc_cc_closed_components(vc_ctx* ctx, vc_cc* c, int64_t* n_closed)
{
    if (check_handle(ctx, c) != VC_OK || !n_closed)
        return VC_ERR_INVALID;
    if (c->nf && !c->fv)
        return vc_fail(ctx, VC_ERR_STATE, "vc_cc_closed_components: no face vertices");
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    *n_closed = 0;
    // every measure 0 and both thresholds 1 (src/highlevelalgo.cpp:1004-1012):
    // everything collapsible goes, closed pockets stay
    vc_ccthin* t = nullptr;
    int rc = vc_ccthin_create(ctx, c, &t);
    if (rc != VC_OK)
        return rc;
    struct Release
    {
        vc_ctx* x;
        vc_ccthin* t;
        ~Release() { vc_ccthin_release(x, t); }
    } release{ctx, t};
    if ((rc = thin_prune(ctx, t, 1.0f, 1.0f, nullptr)) != VC_OK)
        return rc;
    Mem m;
    uint32_t *vpos, *epos, *tpos;
    int64_t NV = 0, NE = 0, NT = 0;
    if ((rc = thin_remaining_pos(ctx, t, m, &vpos, &epos, &tpos, &NV, &NE, &NT)) != VC_OK)
        return rc;
    // only worth counting when faces are left (src/highlevelalgo.cpp:1015-1018)
    if (NT == 0)
        return VC_OK;
    int32_t *dt = nullptr, *par = nullptr;
    uint64_t* keys = nullptr;
    unsigned long long *vals = nullptr, *cnt = nullptr;
    const uint64_t H = table_size(3 * NT);
    auto get = [&](auto** p, size_t n) {
        if (rc == VC_OK)
            rc = m.get(ctx, p, n);
    };
    get(&dt, 3 * (size_t)NT + 3);
    get(&par, (size_t)NT + 1);
    get(&keys, (size_t)H);
    get(&vals, (size_t)H);
    get(&cnt, 1);
    if (rc != VC_OK)
        return rc;
    const int32_t mx = std::max(c->nv, std::max(c->ne, c->nf));
    rem_emit<<<nblk(mx), kB, 0, S>>>(c->nv, c->ne, c->nf, c->E, c->fo, c->fv, vpos, epos,
                                      tpos, nullptr, nullptr, dt, nullptr);
    fill_u64<<<nblk((int64_t)H), kB, 0, S>>>(keys, (int64_t)H, kEmpty);
    fill_u64<<<nblk((int64_t)H), kB, 0, S>>>(reinterpret_cast<uint64_t*>(vals),
                                              (int64_t)H, ~0ull);
    VC_HIP(ctx, hipMemsetAsync(c->err, 0, sizeof(int), S));
    uf_init<<<nblk(NT), kB, 0, S>>>(par, NT);
    tri_sides<<<nblk(3 * NT), kB, 0, S>>>(dt, NT, keys, vals, H - 1);
    tri_union<<<nblk(3 * NT), kB, 0, S>>>(dt, NT, keys, vals, H - 1, par, c->err);
    VC_CHECK_LAUNCH(ctx);
    int herr = 0;
    if ((rc = count_roots(ctx, par, NT, cnt, n_closed)) != VC_OK ||
        (rc = read_err(ctx, c->err, &herr)) != VC_OK)
        return rc;
    if (herr)
        return vc_fail(ctx, VC_ERR_GUARD, "union-find: iteration guard tripped");
    return VC_OK;
}

int vc_euler_char(vc_ctx* ctx, int32_t nv, const int32_t* edges, int64_t ne,
                  const int32_t* tris, int64_t nt, vc_cc_topo* out)
{
    if (!ctx || nv < 0 || ne < 0 || nt < 0 || (ne && !edges) || (nt && !tris) || !out)
        return VC_ERR_INVALID;
    for (int64_t i = 0; i < 2 * ne; ++i)
        if (edges[i] < 0 || edges[i] >= nv)
            return vc_fail(ctx, VC_ERR_INVALID, "vc_euler_char: edge end out of range");
    for (int64_t i = 0; i < 3 * nt; ++i)
        if (tris[i] < 0 || tris[i] >= nv)
            return vc_fail(ctx, VC_ERR_INVALID, "vc_euler_char: corner out of range");
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    Mem m;
    int32_t *de = nullptr, *dt = nullptr;
    int* err = nullptr;
    int rc = m.get(ctx, &de, 2 * (size_t)ne + 2);
    if (rc == VC_OK)
        rc = m.get(ctx, &dt, 3 * (size_t)nt + 3);
    if (rc == VC_OK)
        rc = m.get(ctx, &err, 1);
    if (rc != VC_OK)
        return rc;
    if (ne)
        VC_HIP(ctx, hipMemcpyAsync(de, edges, 8 * (size_t)ne, hipMemcpyHostToDevice, S));
    if (nt)
        VC_HIP(ctx, hipMemcpyAsync(dt, tris, 12 * (size_t)nt, hipMemcpyHostToDevice, S));
    return euler_arrays_dev(ctx, nv, de, ne, dt, nt, err, out);
}

int vc_cc_prune_device(vc_ctx* ctx, const vc_cc_view* cc, const float* em,
                       const float* fm, float f_t, float l_t, uint8_t* rm_v,
                       uint8_t* rm_e, uint8_t* rm_f, int32_t* ref_e,
                       int64_t* generations)
{
    if (!ctx || !cc || cc->nv < 0 || cc->ne < 0 || cc->nf < 0)
        return VC_ERR_INVALID;
    const int32_t nv = cc->nv, ne = cc->ne, nf = cc->nf;
    if ((ne && (!cc->edges || !em || !rm_e || !ref_e)) ||
        (nf && (!cc->face_off || !cc->face_edges || !fm || !rm_f)) ||
        (nv && !rm_v))
        return VC_ERR_INVALID;
    const int64_t nfe = nf ? cc->face_off[nf] : 0;
    if (nfe < 0 || nfe >= (int64_t)UINT32_MAX / 2 ||
        2 * (int64_t)ne >= (int64_t)UINT32_MAX / 2)
        return vc_fail(ctx, VC_ERR_INVALID, "vc_cc_prune_device: complex too large");
    if (generations)
        *generations = 0;
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    // a complex given by its edges and side ids (no face vertices)
    vc_cc c;
    c.ctx = ctx;
    c.nv = nv;
    c.ne = ne;
    c.nf = nf;
    c.nfe = nfe;
    int rc = VC_OK;
    auto get = [&](auto** p, size_t n) {
        if (rc == VC_OK)
            rc = c.mem.get(ctx, p, n);
    };
    get(&c.E, 2 * (size_t)ne + 2);
    get(&c.fo, (size_t)nf + 1);
    get(&c.fe, (size_t)nfe + 1);
    if (rc != VC_OK)
        return rc;
    if (ne)
        VC_HIP(ctx, hipMemcpyAsync(c.E, cc->edges, 8 * (size_t)ne, hipMemcpyHostToDevice, S));
    if (nf)
        VC_HIP(ctx, hipMemcpyAsync(c.fo, cc->face_off, 8 * ((size_t)nf + 1),
                                   hipMemcpyHostToDevice, S));
    if (nfe)
        VC_HIP(ctx, hipMemcpyAsync(c.fe, cc->face_edges, 4 * (size_t)nfe,
                                   hipMemcpyHostToDevice, S));
    if ((rc = cc_adjacency(ctx, &c)) != VC_OK)
        return rc;
    vc_ccthin t;
    t.cc = &c;
    if ((rc = thin_alloc(ctx, &t)) != VC_OK)
        return rc;
    if (ne)
        VC_HIP(ctx, hipMemcpyAsync(t.em, em, 4 * (size_t)ne, hipMemcpyHostToDevice, S));
    if (nf)
        VC_HIP(ctx, hipMemcpyAsync(t.fm, fm, 4 * (size_t)nf, hipMemcpyHostToDevice, S));
    if ((rc = thin_prune(ctx, &t, f_t, l_t, generations)) != VC_OK)
        return rc;
    if (nv)
        VC_HIP(ctx, hipMemcpyAsync(rm_v, t.rm_v, nv, hipMemcpyDeviceToHost, S));
    if (ne)
    {
        VC_HIP(ctx, hipMemcpyAsync(rm_e, t.rm_e, ne, hipMemcpyDeviceToHost, S));
        VC_HIP(ctx, hipMemcpyAsync(ref_e, t.ref_e, 4 * (size_t)ne, hipMemcpyDeviceToHost, S));
    }
    if (nf)
        VC_HIP(ctx, hipMemcpyAsync(rm_f, t.rm_f, nf, hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    return VC_OK;
}

} // extern "C"

// ---- a resident complex from device arrays (vc_internal.h)

int vc_cc_create_dev(vc_ctx* ctx, int32_t nv, int32_t ne_in, const int32_t* dE,
                     int32_t nf, const int64_t* dfo, const int32_t* dfv, int64_t nfv,
                     vc_cc** out)
{
    *out = nullptr;
    if (nfv >= (int64_t)UINT32_MAX / 2 || 2 * ((int64_t)ne_in + nfv) >= (int64_t)INT32_MAX)
        return vc_fail(ctx, VC_ERR_INVALID, "vc_cc_create: complex too large");
    hipStream_t S = ctx->stream;
    vc_cc* c = new (std::nothrow) vc_cc();
    if (!c)
        return VC_ERR_NOMEM;
    c->ctx = ctx;
    c->nv = nv;
    c->nf = nf;
    c->nfe = nfv;
    int rc = VC_OK;
    auto get = [&](Mem& m, auto** p, size_t n) {
        if (rc == VC_OK)
            rc = m.get(ctx, p, n);
    };
    get(c->mem, &c->fo, (size_t)nf + 1);
    get(c->mem, &c->fv, (size_t)nfv + 1);
    get(c->mem, &c->fe, (size_t)nfv + 1);
    get(c->mem, &c->E, 2 * ((size_t)ne_in + nfv) + 2);
    auto fail = [&](int code) {
        (void)hipStreamSynchronize(S);
        delete c;
        return code;
    };
    if (rc != VC_OK)
        return fail(rc);
    hipError_t e = hipSuccess;
    if (nf)
        e = hipMemcpyAsync(c->fo, dfo, 8 * ((size_t)nf + 1), hipMemcpyDeviceToDevice, S);
    if (e == hipSuccess && nfv)
        e = hipMemcpyAsync(c->fv, dfv, 4 * (size_t)nfv, hipMemcpyDeviceToDevice, S);
    if (e != hipSuccess)
        return fail(vc_fail(ctx, VC_ERR_HIP, "vc_cc_create: %s", hipGetErrorString(e)));
    int32_t ne = 0;
    if ((rc = vc_cc_finalize_dev(ctx, ne_in, dE, nf, c->fo, c->fv, nfv, c->E, c->fe,
                                 &ne)) != VC_OK)
        return fail(rc);
    c->ne = ne;
    if ((rc = cc_adjacency(ctx, c)) != VC_OK)
        return fail(rc);
    *out = c;
    return VC_OK;
}

void vc_cc_dview_get(const vc_cc* c, vc_cc_dview* v)
{
    v->nv = c->nv;
    v->ne = c->ne;
    v->nf = c->nf;
    v->nfe = c->nfe;
    v->E = c->E;
    v->fo = c->fo;
    v->fv = c->fv;
    v->fe = c->fe;
    v->ve_off = c->ve_off;
    v->ve = c->ve;
    v->ef_off = c->ef_off;
    v->ef = c->ef;
}
