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
#include <climits>
#include <cstdint>
#include <cstring>
#include <vector>

#include "vc_internal.h"

namespace
{

constexpr int kB = 256;
constexpr int kScanB = 1024;
constexpr int kErrLogic = 1;   // a taken pair with a count other than 1
constexpr int kErrTwice = 2;   // an edge removed twice in one generation
constexpr int kErrInvalid = 3; // out-of-range ids, self-loops, repeated sides

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
    if (a < 0 || b < 0 || a >= nv || b >= nv || a == b)
    {
        atomicMax(err, kErrInvalid);
        return;
    }
    atomicAdd(&cnt[a], 1u);
    atomicAdd(&cnt[b], 1u);
}

// per face: side ids in range, no side listed twice; edge face counts
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
        bool bad = e < 0 || e >= ne;
        for (int64_t k = fo[f]; k < i && !bad; ++k)
            bad = fe[k] == e;
        if (bad)
        {
            atomicMax(err, kErrInvalid);
            return;
        }
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

// ascending ids in every list (insertion sort: lists are short); the list
// length is the element's initial count
__global__ void __launch_bounds__(kB)
    sort_lists(const uint32_t* __restrict__ off, int32_t n,
               int32_t* __restrict__ idx, int32_t* __restrict__ ref)
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
    ref[i] = (int32_t)(hi - lo);
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

__global__ void __launch_bounds__(kB) fill_i32(int32_t* a, int64_t n, int32_t v)
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

// device allocations released at scope end
struct Mem
{
    std::vector<void*> p;
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

} // namespace

extern "C" int vc_cc_prune_device(vc_ctx* ctx, const vc_cc_view* cc,
                                  const float* em, const float* fm, float f_t,
                                  float l_t, uint8_t* rm_v, uint8_t* rm_e,
                                  uint8_t* rm_f, int32_t* ref_e,
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
    Mem m;
    Cx C{};
    int32_t *dE, *dfe, *ve, *ef, *rref_e, *rref_v, *rg_f, *rj_f, *rg_e, *rj_e,
        *first_f, *first_e;
    int64_t* dfo;
    float *dem, *dfm;
    uint32_t *ve_off, *ef_off, *fillv, *fille, *sums;
    uint8_t *drm_v, *drm_e, *drm_f;
    int* err;
    const size_t V = (size_t)nv + 1, E = (size_t)ne + 1, F = (size_t)nf + 1;
    int rc = VC_OK;
    auto get = [&](auto** p, size_t n) {
        if (rc == VC_OK)
            rc = m.get(ctx, p, n);
    };
    get(&dE, 2 * E);
    get(&dfo, F);
    get(&dfe, (size_t)nfe + 1);
    get(&dem, E);
    get(&dfm, F);
    get(&ve_off, V + 1);
    get(&ef_off, E + 1);
    get(&fillv, V);
    get(&fille, E);
    get(&ve, 2 * E);
    get(&ef, (size_t)nfe + 1);
    get(&rref_e, E);
    get(&rref_v, V);
    get(&rg_f, F);
    get(&rj_f, F);
    get(&rg_e, E);
    get(&rj_e, E);
    get(&first_f, F);
    get(&first_e, E);
    get(&drm_v, V);
    get(&drm_e, E);
    get(&drm_f, F);
    get(&sums, (size_t)kScanB + 1);
    get(&err, 4);
    if (rc != VC_OK)
        return rc;
    if (ne)
    {
        VC_HIP(ctx, hipMemcpyAsync(dE, cc->edges, 8 * (size_t)ne,
                                   hipMemcpyHostToDevice, S));
        VC_HIP(ctx, hipMemcpyAsync(dem, em, 4 * (size_t)ne, hipMemcpyHostToDevice, S));
    }
    if (nf)
    {
        VC_HIP(ctx, hipMemcpyAsync(dfo, cc->face_off, 8 * F, hipMemcpyHostToDevice, S));
        VC_HIP(ctx, hipMemcpyAsync(dfm, fm, 4 * (size_t)nf, hipMemcpyHostToDevice, S));
        if (nfe)
            VC_HIP(ctx, hipMemcpyAsync(dfe, cc->face_edges, 4 * (size_t)nfe,
                                       hipMemcpyHostToDevice, S));
    }
    // zero counters / flags with small fill kernels
    VC_HIP(ctx, hipMemsetAsync(ve_off, 0, 4 * (V + 1), S));
    VC_HIP(ctx, hipMemsetAsync(ef_off, 0, 4 * (E + 1), S));
    VC_HIP(ctx, hipMemsetAsync(fillv, 0, 4 * V, S));
    VC_HIP(ctx, hipMemsetAsync(fille, 0, 4 * E, S));
    VC_HIP(ctx, hipMemsetAsync(drm_v, 0, V, S));
    VC_HIP(ctx, hipMemsetAsync(err, 0, sizeof(int), S));
    fill_i32<<<nblk(F), kB, 0, S>>>(rg_f, (int64_t)F, -1);
    fill_i32<<<nblk(E), kB, 0, S>>>(rg_e, (int64_t)E, -1);
    fill_i32<<<nblk(F), kB, 0, S>>>(first_f, (int64_t)F, INT_MAX);
    fill_i32<<<nblk(E), kB, 0, S>>>(first_e, (int64_t)E, INT_MAX);
    VC_CHECK_LAUNCH(ctx);
    if (ne)
        deg_v<<<nblk(ne), kB, 0, S>>>(dE, ne, nv, ve_off, err);
    if (nf)
        deg_e<<<nblk(nf), kB, 0, S>>>(dfo, dfe, nf, ne, ef_off, err);
    VC_CHECK_LAUNCH(ctx);
    int herr = 0;
    VC_HIP(ctx, hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    if (herr)
        return vc_fail(ctx, VC_ERR_INVALID,
                       "vc_cc_prune_device: edge ends out of range or equal, "
                       "or a face side out of range or listed twice");
    if ((rc = scan_any(ctx, ve_off, (int64_t)V, sums)) != VC_OK ||
        (rc = scan_any(ctx, ef_off, (int64_t)E, sums)) != VC_OK)
        return rc;
    if (ne)
        fill_ve<<<nblk(ne), kB, 0, S>>>(dE, ne, ve_off, fillv, ve);
    if (nf)
        fill_ef<<<nblk(nf), kB, 0, S>>>(dfo, dfe, nf, ef_off, fille, ef);
    VC_CHECK_LAUNCH(ctx);
    if (nv)
        sort_lists<<<nblk(nv), kB, 0, S>>>(ve_off, nv, ve, rref_v);
    if (ne)
        sort_lists<<<nblk(ne), kB, 0, S>>>(ef_off, ne, ef, rref_e);
    VC_CHECK_LAUNCH(ctx);

    C.E = dE;
    C.fo = dfo;
    C.fe = dfe;
    C.ve_off = ve_off;
    C.ve = ve;
    C.ef_off = ef_off;
    C.ef = ef;
    C.em = dem;
    C.fm = dfm;
    C.ref_e = rref_e;
    C.ref_v = rref_v;
    C.rg_f = rg_f;
    C.rj_f = rj_f;
    C.rg_e = rg_e;
    C.rj_e = rj_e;
    C.rm_v = drm_v;
    C.first_f = first_f;
    C.first_e = first_e;
    C.err = err;

    // generation 0: face-edge pairs by edge, then edge-vertex pairs by vertex
    const int64_t nq = (int64_t)ne + nv;
    uint32_t* qflag;
    if ((rc = m.get(ctx, &qflag, (size_t)nq + 1)) != VC_OK)
        return rc;
    VC_HIP(ctx, hipMemsetAsync(qflag, 0, 4 * ((size_t)nq + 1), S));
    if (nq)
        init_flags<<<nblk(nq), kB, 0, S>>>(ne, nv, rref_e, rref_v, ef_off, ef,
                                            ve_off, ve, dem, dfm, f_t, l_t, qflag);
    VC_CHECK_LAUNCH(ctx);
    if ((rc = scan_any(ctx, qflag, nq + 1, sums)) != VC_OK)
        return rc;
    uint32_t np = 0;
    VC_HIP(ctx, hipMemcpyAsync(&np, qflag + nq, 4, hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    // pair buffers grow on demand
    size_t cap_p = 0, cap_s = 0;
    Pairs P{}, N{}, SL{};
    uint8_t* eff = nullptr;
    uint32_t *slot_off = nullptr, *sflag = nullptr;
    struct Owned
    {
        std::vector<void*> v;
        void push_back(void* q) { v.push_back(q); }
        ~Owned()
        {
            for (void* q : v)
                vc_free(q);
        }
    } owned;
    auto grow_p = [&](size_t n) -> int {
        if (n <= cap_p)
            return VC_OK;
        size_t c = n + n / 2 + 1024;
        void *t1, *c1, *o1, *t2, *c2, *o2, *e1, *s1;
        int r = VC_OK;
        for (auto* pp : {&t1, &t2, &e1})
            if (r == VC_OK && (r = vc_alloc(ctx, pp, c + 16)) == VC_OK)
                owned.push_back(*pp);
        for (auto* pp : {&c1, &o1, &c2, &o2, &s1})
            if (r == VC_OK && (r = vc_alloc(ctx, pp, 4 * (c + 16))) == VC_OK)
                owned.push_back(*pp);
        if (r != VC_OK)
            return r;
        // keep the current pairs
        if (cap_p)
        {
            VC_HIP(ctx, hipMemcpyAsync(t1, P.t, cap_p, hipMemcpyDeviceToDevice, S));
            VC_HIP(ctx, hipMemcpyAsync(c1, P.c, 4 * cap_p, hipMemcpyDeviceToDevice, S));
            VC_HIP(ctx, hipMemcpyAsync(o1, P.o, 4 * cap_p, hipMemcpyDeviceToDevice, S));
        }
        P = Pairs{(uint8_t*)t1, (int32_t*)c1, (int32_t*)o1};
        N = Pairs{(uint8_t*)t2, (int32_t*)c2, (int32_t*)o2};
        eff = (uint8_t*)e1;
        slot_off = (uint32_t*)s1;
        cap_p = c;
        return VC_OK;
    };
    auto grow_s = [&](size_t n) -> int {
        if (n <= cap_s)
            return VC_OK;
        size_t c = n + n / 2 + 1024;
        void *t, *cc2, *o, *fl;
        int r = VC_OK;
        if ((r = vc_alloc(ctx, &t, c + 16)) == VC_OK)
            owned.push_back(t);
        for (auto* pp : {&cc2, &o, &fl})
            if (r == VC_OK && (r = vc_alloc(ctx, pp, 4 * (c + 16))) == VC_OK)
                owned.push_back(*pp);
        if (r != VC_OK)
            return r;
        SL = Pairs{(uint8_t*)t, (int32_t*)cc2, (int32_t*)o};
        sflag = (uint32_t*)fl;
        cap_s = c;
        return VC_OK;
    };
    if ((rc = grow_p((size_t)np + 1)) != VC_OK)
        return rc;
    if (np)
        init_pairs<<<nblk(nq), kB, 0, S>>>(ne, nv, qflag, qflag + nq, ef_off, ef,
                                            ve_off, ve, P);
    VC_CHECK_LAUNCH(ctx);

    int32_t g = 0;
    while (np > 0 && rc == VC_OK)
    {
        const int32_t n = (int32_t)np;
        first_face<<<nblk(n), kB, 0, S>>>(C, P, n);
        first_edge<<<nblk(n), kB, 0, S>>>(C, P, n, eff, slot_off);
        decide<<<nblk(n), kB, 0, S>>>(C, P, n, eff);
        VC_CHECK_LAUNCH(ctx);
        VC_HIP(ctx, hipMemsetAsync(slot_off + n, 0, 4, S));
        if ((rc = scan_any(ctx, slot_off, (int64_t)n + 1, sums)) != VC_OK)
            break;
        uint32_t hs[2] = {0, 0};
        VC_HIP(ctx, hipMemcpyAsync(&hs[0], slot_off + n, 4, hipMemcpyDeviceToHost, S));
        VC_HIP(ctx, hipMemcpyAsync(&hs[1], err, 4, hipMemcpyDeviceToHost, S));
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
        if ((rc = grow_s((size_t)ns + 1)) != VC_OK)
            break;
        apply<<<nblk(n), kB, 0, S>>>(C, P, n, eff, g);
        fill_u32<<<nblk((int64_t)ns + 1), kB, 0, S>>>(sflag, (int64_t)ns + 1, 0u);
        trigger<<<nblk(n), kB, 0, S>>>(C, P, n, eff, g, slot_off, f_t, l_t, sflag,
                                         SL);
        VC_CHECK_LAUNCH(ctx);
        if ((rc = scan_any(ctx, sflag, (int64_t)ns + 1, sums)) != VC_OK)
            break;
        uint32_t nn = 0;
        VC_HIP(ctx, hipMemcpyAsync(&nn, sflag + ns, 4, hipMemcpyDeviceToHost, S));
        VC_HIP(ctx, hipStreamSynchronize(S));
        if ((rc = grow_p((size_t)nn + 1)) != VC_OK)
            break;
        if (nn)
            compact<<<nblk(ns), kB, 0, S>>>(sflag, ns, sflag + ns, SL, N);
        VC_CHECK_LAUNCH(ctx);
        std::swap(P, N);
        np = nn;
        ++g;
    }
    if (rc == VC_OK)
    {
        const int32_t mx = ne > nf ? ne : nf;
        if (mx)
            results<<<nblk(mx), kB, 0, S>>>(ne, nf, rg_e, rg_f, drm_e, drm_f);
        VC_CHECK_LAUNCH(ctx);
        if (nv)
            VC_HIP(ctx, hipMemcpyAsync(rm_v, drm_v, nv, hipMemcpyDeviceToHost, S));
        if (ne)
        {
            VC_HIP(ctx, hipMemcpyAsync(rm_e, drm_e, ne, hipMemcpyDeviceToHost, S));
            VC_HIP(ctx, hipMemcpyAsync(ref_e, rref_e, 4 * (size_t)ne,
                                       hipMemcpyDeviceToHost, S));
        }
        if (nf)
            VC_HIP(ctx, hipMemcpyAsync(rm_f, drm_f, nf, hipMemcpyDeviceToHost, S));
        VC_HIP(ctx, hipStreamSynchronize(S));
        if (generations)
            *generations = g;
    }
    return rc;
}
