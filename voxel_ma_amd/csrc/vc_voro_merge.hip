// vc_voro_merge.hip -- the close-vertex merge of vc_voro_merge_close on the
// device (VoroInfo::mergeCloseVts, src/voroinfo.cpp:1064-1384).
//
// The merge grows a cluster from every unvisited vertex in order by a BFS
// that absorbs neighbours within eps of the seed.  When no edge joins two
// vertices within eps, no BFS absorbs anything: every (eligible) vertex is
// its own cluster and the merge is a relabelling — old_to_new by a scan,
// faces rewritten with their distinct new ids, the new edge set sorted and
// deduplicated (radix sort of 64-bit pair keys).  That case runs here; the
// others depend on the BFS order and run on the host (vc_voro_merge_close).
#include <hipcub/hipcub.hpp>

#include <cstdint>
#include <type_traits>
#include <vector>

#include "vc_internal.h"
#include "../../include/vc_voro.h"

namespace
{

constexpr int kB = 256;
unsigned nblk(int64_t n) { return (unsigned)((n + kB - 1) / kB); }

// an edge joining two distinct vertices within eps (the host's float
// dist2: separately rounded products and sums)
__global__ void __launch_bounds__(kB)
    close_edges(const float* __restrict__ vts, const int32_t* __restrict__ E,
                int32_t ne, float sq_r, int* any)
{
#pragma clang fp contract(off)
    const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (e >= ne)
        return;
    const int32_t a = E[2 * e], b = E[2 * e + 1];
    if (a == b)
        return;
    const float d0 = vts[3 * b] - vts[3 * a];
    const float d1 = vts[3 * b + 1] - vts[3 * a + 1];
    const float d2 = vts[3 * b + 2] - vts[3 * a + 2];
    float q = d0 * d0;
    q += d1 * d1;
    q += d2 * d2;
    if (q <= sq_r)
        atomicOr(any, 1);
}

__global__ void __launch_bounds__(kB)
    elig(int32_t nv, const uint8_t* __restrict__ valid, int only_inside,
         uint32_t* __restrict__ flag)
{
    const int64_t v = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (v < nv)
        flag[v] = (!only_inside || valid[v]) ? 1u : 0u;
}

__global__ void __launch_bounds__(kB)
    relabel(int32_t nv, const uint32_t* __restrict__ pos,
            int32_t* __restrict__ o2n, int32_t* __restrict__ new_vts)
{
    const int64_t v = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (v >= nv)
        return;
    const bool e = pos[v + 1] != pos[v];
    o2n[v] = e ? (int32_t)pos[v] : -1;
    if (e)
        new_vts[pos[v]] = (int32_t)v;
}

__device__ __forceinline__ uint64_t ekey(int32_t u, int32_t w)
{
    if (u > w)
    {
        const int32_t t = u;
        u = w;
        w = t;
    }
    return ((uint64_t)(uint32_t)u << 32) | (uint32_t)w;
}

// kept edges: key slot e (flag 1)
__global__ void __launch_bounds__(kB)
    edge_keys(const int32_t* __restrict__ E, int32_t ne,
              const uint8_t* __restrict__ valid, int only_inside,
              const int32_t* __restrict__ o2n, uint8_t* __restrict__ keep,
              uint64_t* __restrict__ key, uint32_t* __restrict__ kflag)
{
    const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (e >= ne)
        return;
    const int32_t a = E[2 * e], b = E[2 * e + 1];
    bool k = !(only_inside && !(valid[a] && valid[b]));
    const int32_t u = o2n[a], w = o2n[b];
    k = k && u != w;
    keep[e] = k ? 1 : 0;
    kflag[e] = k ? 1u : 0u;
    key[e] = ekey(u, w);
}

// per face: removed (invalid, < 2 distinct new ids, or 2: an edge key in
// slot ne + f) or kept with k distinct ids (first-occurrence order)
__global__ void __launch_bounds__(kB)
    face_pass(const int64_t* __restrict__ fo, const int32_t* __restrict__ fv,
              int32_t nf, int32_t ne, const uint8_t* __restrict__ valid,
              int only_inside, const int32_t* __restrict__ o2n,
              uint8_t* __restrict__ remove, uint32_t* __restrict__ fsize,
              uint32_t* __restrict__ fkeep, uint64_t* __restrict__ key,
              uint32_t* __restrict__ kflag)
{
    const int64_t f = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (f >= nf)
        return;
    bool ok = true;
    if (only_inside)
        for (int64_t j = fo[f]; j < fo[f + 1] && ok; ++j)
            ok = valid[fv[j]] != 0;
    uint32_t k = 0;
    int32_t c0 = -1, c1 = -1;
    if (ok)
        for (int64_t j = fo[f]; j < fo[f + 1]; ++j)
        {
            const int32_t c = o2n[fv[j]];
            bool seen = false;
            for (int64_t i = fo[f]; i < j && !seen; ++i)
                seen = o2n[fv[i]] == c;
            if (seen)
                continue;
            if (k == 0)
                c0 = c;
            else if (k == 1)
                c1 = c;
            ++k;
        }
    const bool kept = ok && k >= 3;
    remove[f] = kept ? 0 : 1;
    fsize[f] = kept ? k : 0u;
    fkeep[f] = kept ? 1u : 0u;
    kflag[ne + f] = ok && k == 2 ? 1u : 0u;
    key[ne + f] = ekey(c0, c1);
}

__global__ void __launch_bounds__(kB)
    face_write(const int64_t* __restrict__ fo, const int32_t* __restrict__ fv,
               int32_t nf, const int32_t* __restrict__ o2n,
               const uint32_t* __restrict__ fpos, const uint32_t* __restrict__ kpos,
               int64_t* __restrict__ nfo, int32_t* __restrict__ nfv)
{
    const int64_t f = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (f >= nf || kpos[f + 1] == kpos[f])
        return;
    int64_t o = fpos[f];
    for (int64_t j = fo[f]; j < fo[f + 1]; ++j)
    {
        const int32_t c = o2n[fv[j]];
        bool seen = false;
        for (int64_t i = fo[f]; i < j && !seen; ++i)
            seen = o2n[fv[i]] == c;
        if (!seen)
            nfv[o++] = c;
    }
    nfo[kpos[f] + 1] = o;
}

__global__ void __launch_bounds__(kB)
    gather_keys(int64_t n, const uint64_t* __restrict__ key,
                const uint32_t* __restrict__ pos, uint64_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i < n && pos[i + 1] != pos[i])
        out[pos[i]] = key[i];
}

__global__ void __launch_bounds__(kB)
    uniq_flags(int64_t n, const uint64_t* __restrict__ k, uint32_t* __restrict__ f)
{
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i < n)
        f[i] = (i == 0 || k[i] != k[i - 1]) ? 1u : 0u;
}

__global__ void __launch_bounds__(kB)
    uniq_write(int64_t n, const uint64_t* __restrict__ k,
               const uint32_t* __restrict__ pos, int32_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i < n && pos[i + 1] != pos[i])
    {
        out[2 * (int64_t)pos[i]] = (int32_t)(k[i] >> 32);
        out[2 * (int64_t)pos[i] + 1] = (int32_t)(k[i] & 0xFFFFFFFFull);
    }
}

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

} // namespace

void vc_voro_mg_free(vc_voro_mg* m)
{
    vc_free(m->o2n);
    vc_free(m->nvts);
    vc_free(m->nfo);
    vc_free(m->nfv);
    vc_free(m->keep);
    vc_free(m->rem);
    vc_free(m->fpos);
    vc_free(m->edges);
    *m = vc_voro_mg{};
}

int vc_voro_merge_dev(vc_ctx* ctx, int32_t nv, int32_t ne, int32_t nf, const float* dv,
                      const int32_t* dE, const int64_t* dfo, const int32_t* dfv,
                      int64_t nfv, const uint8_t* dval, float eps, int only_inside,
                      vc_voro_mg* out, int* merges)
{
    *out = vc_voro_mg{};
    *merges = 0;
    hipStream_t S = ctx->stream;
    Mem m;
    int* any;
    uint32_t *vflag, *kflag, *fsize, *upos;
    uint64_t *key, *kin, *kout;
    int rc = VC_OK;
    auto get = [&](auto** p, size_t n) {
        if (rc == VC_OK)
            rc = m.get(ctx, p, n);
    };
    auto own = [&](auto** p, size_t n) {
        if (rc == VC_OK)
        {
            void* q = nullptr;
            rc = vc_alloc(ctx, &q, n * sizeof(**p) + 16);
            *p = static_cast<std::remove_reference_t<decltype(*p)>>(q);
        }
    };
    get(&any, 1);
    if (rc != VC_OK)
        return rc;
    VC_HIP(ctx, hipMemsetAsync(any, 0, sizeof(int), S));
    if (ne)
        close_edges<<<nblk(ne), kB, 0, S>>>(dv, dE, ne, eps * eps, any);
    VC_CHECK_LAUNCH(ctx);
    int hany = 0;
    VC_HIP(ctx, hipMemcpyAsync(&hany, any, sizeof(int), hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    if (hany)
    {
        // some BFS absorbs a neighbour: the merge depends on the visit order
        *merges = 1;
        return VC_OK;
    }
    const size_t NK = (size_t)ne + nf;
    own(&out->o2n, (size_t)nv + 1);
    own(&out->nvts, (size_t)nv + 1);
    own(&out->nfo, (size_t)nf + 2);
    own(&out->nfv, (size_t)nfv + 1);
    own(&out->keep, (size_t)ne + 1);
    own(&out->rem, (size_t)nf + 1);
    own(&out->fpos, (size_t)nf + 1);
    get(&vflag, (size_t)nv + 1);
    get(&kflag, NK + 1);
    get(&fsize, (size_t)nf + 1);
    get(&key, NK + 1);
    get(&kin, NK + 1);
    get(&kout, NK + 1);
    get(&upos, NK + 1);
    auto fail = [&](int code) {
        (void)hipStreamSynchronize(S);
        vc_voro_mg_free(out);
        return code;
    };
    if (rc != VC_OK)
        return fail(rc);
    hipError_t e = hipSuccess;
    auto hip = [&](hipError_t r) {
        if (e == hipSuccess)
            e = r;
    };
    hip(hipMemsetAsync(vflag, 0, 4 * ((size_t)nv + 1), S));
    hip(hipMemsetAsync(kflag, 0, 4 * (NK + 1), S));
    hip(hipMemsetAsync(fsize, 0, 4 * ((size_t)nf + 1), S));
    hip(hipMemsetAsync(out->fpos, 0, 4 * ((size_t)nf + 1), S));
    hip(hipMemsetAsync(out->nfo, 0, 8, S));
    if (nv)
        elig<<<nblk(nv), kB, 0, S>>>(nv, dval, only_inside, vflag);
    hip(hipGetLastError());
    if (e != hipSuccess)
        return fail(vc_fail(ctx, VC_ERR_HIP, "merge: %s", hipGetErrorString(e)));
    if ((rc = vc_scan_u32(ctx, vflag, (int64_t)nv + 1)) != VC_OK)
        return fail(rc);
    if (nv)
        relabel<<<nblk(nv), kB, 0, S>>>(nv, vflag, out->o2n, out->nvts);
    if (ne)
        edge_keys<<<nblk(ne), kB, 0, S>>>(dE, ne, dval, only_inside, out->o2n, out->keep,
                                           key, kflag);
    if (nf)
        face_pass<<<nblk(nf), kB, 0, S>>>(dfo, dfv, nf, ne, dval, only_inside, out->o2n,
                                           out->rem, fsize, out->fpos, key, kflag);
    hip(hipGetLastError());
    if (e != hipSuccess)
        return fail(vc_fail(ctx, VC_ERR_HIP, "merge: %s", hipGetErrorString(e)));
    if ((rc = vc_scan_u32(ctx, fsize, (int64_t)nf + 1)) != VC_OK ||
        (rc = vc_scan_u32(ctx, out->fpos, (int64_t)nf + 1)) != VC_OK ||
        (rc = vc_scan_u32(ctx, kflag, (int64_t)NK + 1)) != VC_OK)
        return fail(rc);
    if (nf)
        face_write<<<nblk(nf), kB, 0, S>>>(dfo, dfv, nf, out->o2n, fsize, out->fpos,
                                           out->nfo, out->nfv);
    if (NK)
        gather_keys<<<nblk((int64_t)NK), kB, 0, S>>>((int64_t)NK, key, kflag, kin);
    hip(hipGetLastError());
    uint32_t cnt[3] = {0, 0, 0}; // new vertices, kept faces, keys
    hip(hipMemcpy(&cnt[0], vflag + nv, 4, hipMemcpyDeviceToHost));
    hip(hipMemcpy(&cnt[1], out->fpos + nf, 4, hipMemcpyDeviceToHost));
    hip(hipMemcpy(&cnt[2], kflag + NK, 4, hipMemcpyDeviceToHost));
    int64_t nfv_new = 0;
    hip(hipMemcpy(&nfv_new, out->nfo + cnt[1], 8, hipMemcpyDeviceToHost));
    if (e != hipSuccess)
        return fail(vc_fail(ctx, VC_ERR_HIP, "merge: %s", hipGetErrorString(e)));
    const int64_t nk = cnt[2];
    int64_t nuniq = 0;
    if (nk > 0)
    {
        size_t tmp_bytes = 0;
        hip(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, kin, kout, (int)nk, 0,
                                              64, S));
        void* tmp = nullptr;
        if (e == hipSuccess &&
            (rc = m.get(ctx, reinterpret_cast<uint8_t**>(&tmp), tmp_bytes)) != VC_OK)
            return fail(rc);
        hip(hipcub::DeviceRadixSort::SortKeys(tmp, tmp_bytes, kin, kout, (int)nk, 0, 64,
                                              S));
        hip(hipMemsetAsync(upos + nk, 0, 4, S));
        if (e == hipSuccess)
            uniq_flags<<<nblk(nk), kB, 0, S>>>(nk, kout, upos);
        hip(hipGetLastError());
        if (e != hipSuccess)
            return fail(vc_fail(ctx, VC_ERR_HIP, "merge: %s", hipGetErrorString(e)));
        if ((rc = vc_scan_u32(ctx, upos, nk + 1)) != VC_OK)
            return fail(rc);
        uint32_t u = 0;
        hip(hipMemcpy(&u, upos + nk, 4, hipMemcpyDeviceToHost));
        nuniq = u;
    }
    own(&out->edges, 2 * (size_t)nuniq + 2);
    if (rc != VC_OK)
        return fail(rc);
    if (nuniq && e == hipSuccess)
        uniq_write<<<nblk(nk), kB, 0, S>>>(nk, kout, upos, out->edges);
    hip(hipGetLastError());
    // the scratch is freed on return
    hip(hipStreamSynchronize(S));
    if (e != hipSuccess)
        return fail(vc_fail(ctx, VC_ERR_HIP, "merge: %s", hipGetErrorString(e)));
    out->n_new = (int32_t)cnt[0];
    out->nf_kept = (int32_t)cnt[1];
    out->nfv_new = nfv_new;
    out->ne_new = nuniq;
    return VC_OK;
}

extern "C" int vc_voro_merge_close_device(
    vc_ctx* ctx, int32_t nv, int32_t ne, int32_t nf, const float* vts,
    const int32_t* edges, const int64_t* face_off, const int32_t* face_vts,
    const int64_t* ve_off, const int32_t* ve, const uint8_t* v_valid, float eps,
    int only_inside, int32_t* old_to_new, int32_t* new_vts, int32_t* n_new_vts,
    int64_t* new_face_off, int32_t* new_face_vts, uint8_t* remove_face,
    int32_t* new_edges, int64_t edge_cap, int64_t* n_new_edges,
    uint8_t* keep_edge, int* on_device)
{
    if (on_device)
        *on_device = 0;
    if (!ctx || nv < 0 || ne < 0 || nf < 0 || (nv > 0 && (!vts || !old_to_new || !new_vts)) ||
        (ne > 0 && (!edges || !keep_edge)) ||
        (nf > 0 && (!face_off || !face_vts || !new_face_off || !new_face_vts ||
                    !remove_face)) ||
        (only_inside && nv > 0 && !v_valid) || !n_new_vts || !n_new_edges ||
        !(eps >= 0.0f))
        return VC_ERR_INVALID;
    const int64_t nfv = nf > 0 ? face_off[nf] : 0;
    if ((nf > 0 && face_off[0] != 0) || nfv < 0 || nfv >= (int64_t)INT32_MAX)
        return VC_ERR_INVALID;
    for (int64_t i = 0; i < 2LL * ne; ++i)
        if (edges[i] < 0 || edges[i] >= nv)
            return VC_ERR_INVALID;
    for (int32_t f = 0; f < nf; ++f)
        if (face_off[f + 1] < face_off[f])
            return VC_ERR_INVALID;
    for (int64_t j = 0; j < nfv; ++j)
        if (face_vts[j] < 0 || face_vts[j] >= nv)
            return VC_ERR_INVALID;
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    Mem m;
    float* dv;
    int32_t *dE, *dfv;
    int64_t* dfo;
    uint8_t* dval;
    int rc = VC_OK;
    auto get = [&](auto** p, size_t n) {
        if (rc == VC_OK)
            rc = m.get(ctx, p, n);
    };
    get(&dv, 3 * (size_t)nv + 3);
    get(&dE, 2 * (size_t)ne + 2);
    get(&dfo, (size_t)nf + 1);
    get(&dfv, (size_t)nfv + 1);
    get(&dval, (size_t)nv + 1);
    if (rc != VC_OK)
        return rc;
    if (nv)
        VC_HIP(ctx, hipMemcpyAsync(dv, vts, 12 * (size_t)nv, hipMemcpyHostToDevice, S));
    if (ne)
        VC_HIP(ctx, hipMemcpyAsync(dE, edges, 8 * (size_t)ne, hipMemcpyHostToDevice, S));
    if (nf)
    {
        VC_HIP(ctx, hipMemcpyAsync(dfo, face_off, 8 * ((size_t)nf + 1),
                                   hipMemcpyHostToDevice, S));
        if (nfv)
            VC_HIP(ctx, hipMemcpyAsync(dfv, face_vts, 4 * (size_t)nfv,
                                       hipMemcpyHostToDevice, S));
    }
    if (only_inside && nv)
        VC_HIP(ctx, hipMemcpyAsync(dval, v_valid, (size_t)nv, hipMemcpyHostToDevice, S));
    vc_voro_mg g;
    struct Free
    {
        vc_voro_mg* g;
        ~Free() { vc_voro_mg_free(g); }
    } fr{&g};
    int merges = 0;
    if ((rc = vc_voro_merge_dev(ctx, nv, ne, nf, dv, dE, dfo, dfv, nfv, dval, eps,
                                only_inside, &g, &merges)) != VC_OK)
        return rc;
    if (merges)
        // some BFS absorbs a neighbour: the merge depends on the visit order
        return vc_voro_merge_close(nv, ne, nf, vts, edges, face_off, face_vts,
                                   ve_off, ve, v_valid, eps, only_inside,
                                   old_to_new, new_vts, n_new_vts, new_face_off,
                                   new_face_vts, remove_face, new_edges,
                                   edge_cap, n_new_edges, keep_edge);
    *n_new_vts = g.n_new;
    *n_new_edges = g.ne_new;
    if (nv)
    {
        VC_HIP(ctx, hipMemcpyAsync(old_to_new, g.o2n, 4 * (size_t)nv,
                                   hipMemcpyDeviceToHost, S));
        if (g.n_new)
            VC_HIP(ctx, hipMemcpyAsync(new_vts, g.nvts, 4 * (size_t)g.n_new,
                                       hipMemcpyDeviceToHost, S));
    }
    if (ne)
        VC_HIP(ctx, hipMemcpyAsync(keep_edge, g.keep, (size_t)ne, hipMemcpyDeviceToHost, S));
    if (nf)
    {
        VC_HIP(ctx, hipMemcpyAsync(remove_face, g.rem, (size_t)nf, hipMemcpyDeviceToHost, S));
        VC_HIP(ctx, hipMemcpyAsync(new_face_off, g.nfo, 8 * ((size_t)g.nf_kept + 1),
                                   hipMemcpyDeviceToHost, S));
        if (g.nfv_new)
            VC_HIP(ctx, hipMemcpyAsync(new_face_vts, g.nfv, 4 * (size_t)g.nfv_new,
                                       hipMemcpyDeviceToHost, S));
    }
    if (new_edges && g.ne_new)
    {
        if (edge_cap < g.ne_new)
        {
            VC_HIP(ctx, hipStreamSynchronize(S));
            return VC_ERR_CAPACITY;
        }
        VC_HIP(ctx, hipMemcpyAsync(new_edges, g.edges, 8 * (size_t)g.ne_new,
                                   hipMemcpyDeviceToHost, S));
    }
    VC_HIP(ctx, hipStreamSynchronize(S));
    if (on_device)
        *on_device = 1;
    return VC_OK;
}
