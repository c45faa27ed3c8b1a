// vc_vdiag.hip -- the discrete Voronoi diagram kept resident on the device
// through the voxel-core steps of vol2ma (include/vc_vdiag.h), gfx950.
//
// The diagram lives in device arrays (vertex coordinates, face sites, site
// positions, vertex validity and radii) plus a resident finalized complex
// (vc_cc: edges, face polygons and sides, vertex->edge and edge->face
// adjacency).  Every step runs the same kernels as its host-array entry point
// in include/vc_voro.h, on these arrays; a merge replaces them by the merged
// diagram.
#include <cstring>
#include <new>

#include "vc_internal.h"
#include "../../include/vc_cc.h"
#include "../../include/vc_vdiag.h"

struct vc_vdiag
{
    vc_ctx* ctx = nullptr;
    int32_t nv = 0, nf = 0, ns = 0;
    float* vts = nullptr;    // 3 nv
    int32_t* fs = nullptr;   // 2 nf
    float* sites = nullptr;  // 3 ns
    uint8_t* valid = nullptr; // nv
    float* radii = nullptr;  // nv, once computed
    vc_cc* cc = nullptr;     // the finalized complex
    bool tagged = false;
    int64_t n_valid = 0;
    float bb[6] = {0, 0, 0, 0, 0, 0};
};

namespace
{

constexpr int kB = 256;
unsigned nblk(int64_t n) { return (unsigned)((n + kB - 1) / kB); }

template <typename T>
int dalloc(vc_ctx* ctx, T** p, size_t n)
{
    void* q = nullptr;
    const int rc = vc_alloc(ctx, &q, n * sizeof(T) + 16);
    *p = static_cast<T*>(q);
    return rc;
}

void free_arrays(vc_vdiag* d)
{
    vc_free(d->vts);
    vc_free(d->fs);
    vc_free(d->sites);
    vc_free(d->valid);
    vc_free(d->radii);
    d->vts = nullptr;
    d->fs = nullptr;
    d->sites = nullptr;
    d->valid = nullptr;
    d->radii = nullptr;
    if (d->cc)
        vc_cc_release(d->ctx, d->cc);
    d->cc = nullptr;
}

int check(vc_ctx* ctx, const vc_vdiag* d)
{
    return (ctx && d && d->ctx == ctx) ? VC_OK : VC_ERR_INVALID;
}

__global__ void __launch_bounds__(kB) offsets_step(int64_t n, int64_t step, int64_t* o)
{
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i <= n)
        o[i] = i * step;
}

// float -> unsigned with the same order (no NaN among vertex coordinates)
__device__ __forceinline__ uint32_t ord(float f)
{
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
inline float unord(uint32_t u)
{
    const uint32_t b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

// count of valid vertices and their box: acc[0] count, acc[1..3] ordered
// minima, acc[4..6] ordered maxima
__global__ void __launch_bounds__(kB)
    valid_box(const float* __restrict__ vts, const uint8_t* __restrict__ valid,
              int32_t nv, unsigned long long* cnt, uint32_t* box)
{
    // wavefront reductions first: one add and six min / max per wavefront
    const int32_t v = blockIdx.x * kB + threadIdx.x;
    const bool ok = v < nv && valid[v];
    const unsigned long long c = __popcll(__ballot(ok));
    uint32_t lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        const uint32_t u = ok ? ord(vts[3 * v + a]) : 0u;
        lo[a] = ok ? u : ~0u;
        hi[a] = ok ? u : 0u;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1)
        {
            lo[a] = min(lo[a], (uint32_t)__shfl_xor(lo[a], d));
            hi[a] = max(hi[a], (uint32_t)__shfl_xor(hi[a], d));
        }
    }
    if ((threadIdx.x & 63) != 0 || !c)
        return;
    atomicAdd(cnt, c);
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        atomicMin(&box[a], lo[a]);
        atomicMax(&box[3 + a], hi[a]);
    }
}

__global__ void __launch_bounds__(kB)
    gather_seeds(int32_t n, const int32_t* __restrict__ seeds,
                 const float* __restrict__ vts, const uint8_t* __restrict__ valid,
                 const float* __restrict__ radii, float* __restrict__ nvts,
                 uint8_t* __restrict__ nvalid, float* __restrict__ nradii)
{
    const int32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n)
        return;
    const int32_t s = seeds[i];
    nvts[3 * i] = vts[3 * s];
    nvts[3 * i + 1] = vts[3 * s + 1];
    nvts[3 * i + 2] = vts[3 * s + 2];
    nvalid[i] = valid[s];
    if (radii)
        nradii[i] = radii[s];
}

__global__ void __launch_bounds__(kB)
    gather_face_sites(int32_t nf, const uint8_t* __restrict__ rem,
                      const uint32_t* __restrict__ fpos,
                      const int32_t* __restrict__ fs, int32_t* __restrict__ nfs)
{
    const int32_t f = blockIdx.x * kB + threadIdx.x;
    if (f >= nf || rem[f])
        return;
    nfs[2 * (int64_t)fpos[f]] = fs[2 * f];
    nfs[2 * (int64_t)fpos[f] + 1] = fs[2 * f + 1];
}

// vertex validity of the extraction: the tag (or every vertex) and, for the
// measures, a face is valid iff all its vertices are
__global__ void __launch_bounds__(kB)
    eff_valid(int32_t nv, const uint8_t* __restrict__ valid, int all,
              uint8_t* __restrict__ out)
{
    const int32_t v = blockIdx.x * kB + threadIdx.x;
    if (v < nv)
        out[v] = (all || valid[v]) ? 1 : 0;
}

__global__ void __launch_bounds__(kB)
    face_valid(int32_t nf, const int64_t* __restrict__ fo, const int32_t* __restrict__ fv,
               const uint8_t* __restrict__ vval, uint8_t* __restrict__ out)
{
    const int32_t f = blockIdx.x * kB + threadIdx.x;
    if (f >= nf)
        return;
    uint8_t ok = 1;
    for (int64_t j = fo[f]; j < fo[f + 1]; ++j)
        ok &= vval[fv[j]];
    out[f] = ok;
}

__global__ void __launch_bounds__(kB)
    gather_vm(int64_t n, const int32_t* __restrict__ vi, const float* __restrict__ lv,
              float* __restrict__ vm)
{
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i < n)
        vm[i] = lv[vi[i]];
}

// device scratch freed on every exit path
struct Tmp
{
    void* p[16] = {};
    int n = 0;
    ~Tmp()
    {
        for (int i = 0; i < n; ++i)
            vc_free(p[i]);
    }
    template <typename T>
    int get(vc_ctx* ctx, T** out, size_t count)
    {
        const int rc = dalloc(ctx, out, count);
        if (rc == VC_OK)
            p[n++] = *out;
        return rc;
    }
};

} // namespace

extern "C" {

int vc_vdiag_from_ft(vc_ctx* ctx, vc_vol* vol, vc_vdiag** out)
{
    if (!ctx || !vol || vol->ctx != ctx || !out)
        return VC_ERR_INVALID;
    *out = nullptr;
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    vc_voro_dev g;
    int rc = vc_voro_from_ft_dev(ctx, vol, &g);
    if (rc != VC_OK)
        return rc;
    if (g.nv >= (int64_t)INT32_MAX || g.ns >= (int64_t)INT32_MAX)
    {
        vc_voro_dev_free(&g);
        return vc_fail(ctx, VC_ERR_INVALID, "vc_vdiag_from_ft: diagram too large");
    }
    vc_vdiag* d = new (std::nothrow) vc_vdiag();
    if (!d)
    {
        vc_voro_dev_free(&g);
        return VC_ERR_NOMEM;
    }
    d->ctx = ctx;
    d->nv = (int32_t)g.nv;
    d->nf = (int32_t)g.nf;
    d->ns = (int32_t)g.ns;
    d->vts = g.vts;
    d->fs = g.fs;
    d->sites = g.sites;
    int32_t* quads = g.faces;
    {
        Tmp t;
        int64_t* fo = nullptr;
        rc = t.get(ctx, &fo, (size_t)d->nf + 1);
        if (rc == VC_OK)
            rc = dalloc(ctx, &d->valid, (size_t)d->nv + 1);
        if (rc == VC_OK)
        {
            offsets_step<<<nblk((int64_t)d->nf + 1), kB, 0, S>>>(d->nf, 4, fo);
            hipError_t e = hipGetLastError();
            if (e == hipSuccess)
                e = hipMemsetAsync(d->valid, 0, (size_t)d->nv + 1, S);
            if (e != hipSuccess)
                rc = vc_fail(ctx, VC_ERR_HIP, "vc_vdiag_from_ft: %s", hipGetErrorString(e));
        }
        if (rc == VC_OK)
            rc = vc_cc_create_dev(ctx, d->nv, 0, nullptr, d->nf, fo, quads, 4 * (int64_t)d->nf,
                                  &d->cc);
        (void)hipStreamSynchronize(S);
    }
    vc_free(quads);
    if (rc != VC_OK)
    {
        free_arrays(d);
        delete d;
        return rc;
    }
    *out = d;
    return VC_OK;
}

int vc_vdiag_release(vc_ctx* ctx, vc_vdiag* d)
{
    if (check(ctx, d) != VC_OK)
        return VC_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    free_arrays(d);
    delete d;
    return VC_OK;
}

int vc_vdiag_info_get(const vc_vdiag* d, vc_vdiag_info* info)
{
    if (!d || !info)
        return VC_ERR_INVALID;
    vc_cc_dview v;
    vc_cc_dview_get(d->cc, &v);
    std::memset(info, 0, sizeof(*info));
    info->nv = d->nv;
    info->ne = v.ne;
    info->nf = d->nf;
    info->ns = d->ns;
    info->n_sides = v.nfe;
    info->tagged = d->tagged ? 1 : 0;
    info->radii = d->radii ? 1 : 0;
    info->n_valid = d->n_valid;
    for (int a = 0; a < 3; ++a)
    {
        info->bb_min[a] = d->bb[a];
        info->bb_max[a] = d->bb[3 + a];
    }
    return VC_OK;
}

int vc_vdiag_radii(vc_ctx* ctx, vc_vdiag* d)
{
    if (check(ctx, d) != VC_OK)
        return VC_ERR_INVALID;
    VC_HIP(ctx, hipSetDevice(ctx->device));
    if (!d->radii)
    {
        const int rc = dalloc(ctx, &d->radii, (size_t)d->nv + 1);
        if (rc != VC_OK)
            return rc;
    }
    vc_cc_dview v;
    vc_cc_dview_get(d->cc, &v);
    Tmp t;
    int32_t* last = nullptr;
    int rc = t.get(ctx, &last, (size_t)d->nv + 1);
    if (rc == VC_OK)
        rc = vc_voro_radii_dev(ctx, d->nv, d->nf, d->vts, v.fo, v.fv, d->sites, d->fs,
                               last, d->radii);
    VC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return rc;
}

int vc_vdiag_tag(vc_ctx* ctx, vc_vdiag* d, vc_vol* vol)
{
    if (check(ctx, d) != VC_OK || !vol || vol->ctx != ctx)
        return VC_ERR_INVALID;
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    Tmp t;
    unsigned long long* cnt = nullptr;
    uint32_t* box = nullptr;
    int rc = t.get(ctx, &cnt, 1);
    if (rc == VC_OK)
        rc = t.get(ctx, &box, 6);
    if (rc != VC_OK)
        return rc;
    const uint32_t init[6] = {~0u, ~0u, ~0u, 0u, 0u, 0u};
    VC_HIP(ctx, hipMemsetAsync(cnt, 0, 8, S));
    VC_HIP(ctx, hipMemcpyAsync(box, init, sizeof(init), hipMemcpyHostToDevice, S));
    if ((rc = vc_voro_tag_dev(ctx, vol, d->vts, d->nv, d->valid)) != VC_OK)
        return rc;
    if (d->nv)
        valid_box<<<nblk(d->nv), kB, 0, S>>>(d->vts, d->valid, d->nv, cnt, box);
    VC_CHECK_LAUNCH(ctx);
    unsigned long long hc = 0;
    uint32_t hb[6];
    VC_HIP(ctx, hipMemcpyAsync(&hc, cnt, 8, hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipMemcpyAsync(hb, box, sizeof(hb), hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    d->tagged = true;
    d->n_valid = (int64_t)hc;
    for (int a = 0; a < 6; ++a)
        d->bb[a] = hc ? unord(hb[a]) : 0.0f;
    return VC_OK;
}

int vc_vdiag_merge_close(vc_ctx* ctx, vc_vdiag* d, float eps, int only_inside,
                         int* on_device)
{
    if (check(ctx, d) != VC_OK || !on_device || !(eps >= 0.0f))
        return VC_ERR_INVALID;
    *on_device = 0;
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    vc_cc_dview v;
    vc_cc_dview_get(d->cc, &v);
    vc_voro_mg g;
    struct FreeMg
    {
        vc_voro_mg* g;
        ~FreeMg() { vc_voro_mg_free(g); }
    } fr{&g};
    int merges = 0;
    int rc = vc_voro_merge_dev(ctx, d->nv, v.ne, d->nf, d->vts, v.E, v.fo, v.fv, v.nfe,
                               d->valid, eps, only_inside, &g, &merges);
    if (rc != VC_OK || merges)
        return rc;
    // the merged diagram: seeds with their validity and radius, kept faces
    // with their sites, finalized again over the sorted new edges
    vc_vdiag n;
    n.ctx = ctx;
    n.nv = g.n_new;
    n.nf = g.nf_kept;
    n.ns = d->ns;
    n.sites = d->sites;
    auto drop = [&](int code) {
        (void)hipStreamSynchronize(S);
        n.sites = nullptr; // still d's
        free_arrays(&n);
        return code;
    };
    if ((rc = dalloc(ctx, &n.vts, 3 * (size_t)n.nv + 3)) != VC_OK ||
        (rc = dalloc(ctx, &n.valid, (size_t)n.nv + 1)) != VC_OK ||
        (rc = dalloc(ctx, &n.fs, 2 * (size_t)n.nf + 2)) != VC_OK ||
        (d->radii && (rc = dalloc(ctx, &n.radii, (size_t)n.nv + 1)) != VC_OK))
        return drop(rc);
    if (n.nv)
        gather_seeds<<<nblk(n.nv), kB, 0, S>>>(n.nv, g.nvts, d->vts, d->valid, d->radii,
                                               n.vts, n.valid, n.radii);
    if (d->nf)
        gather_face_sites<<<nblk(d->nf), kB, 0, S>>>(d->nf, g.rem, g.fpos, d->fs, n.fs);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return drop(vc_fail(ctx, VC_ERR_HIP, "vc_vdiag_merge_close: %s", hipGetErrorString(e)));
    if ((rc = vc_cc_create_dev(ctx, n.nv, (int32_t)g.ne_new, g.edges, n.nf, g.nfo, g.nfv,
                               g.nfv_new, &n.cc)) != VC_OK)
        return drop(rc);
    VC_HIP(ctx, hipStreamSynchronize(S));
    // every seed is valid when only the valid vertices seed; otherwise every
    // vertex is its own seed and the flags are unchanged
    const int64_t n_valid = only_inside ? (int64_t)n.nv : d->n_valid;
    d->sites = nullptr; // moves to the merged diagram
    free_arrays(d);
    d->nv = n.nv;
    d->nf = n.nf;
    d->vts = n.vts;
    d->fs = n.fs;
    d->sites = n.sites;
    d->valid = n.valid;
    d->radii = n.radii;
    d->cc = n.cc;
    d->n_valid = n_valid;
    n.vts = nullptr;
    n.fs = nullptr;
    n.sites = nullptr;
    n.valid = nullptr;
    n.radii = nullptr;
    n.cc = nullptr;
    *on_device = 1;
    return VC_OK;
}

int vc_vdiag_get(vc_ctx* ctx, vc_vdiag* d, float* vts, int32_t* edges,
                 int64_t* face_off, int32_t* face_vts, int32_t* face_sites,
                 float* sites, float* radii, uint8_t* valid)
{
    if (check(ctx, d) != VC_OK)
        return VC_ERR_INVALID;
    if (radii && !d->radii)
        return vc_fail(ctx, VC_ERR_STATE, "vc_vdiag_get: radii not computed");
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    vc_cc_dview v;
    vc_cc_dview_get(d->cc, &v);
    auto down = [&](void* h, const void* dp, size_t bytes) {
        return (h && bytes) ? hipMemcpyAsync(h, dp, bytes, hipMemcpyDeviceToHost, S)
                            : hipSuccess;
    };
    VC_HIP(ctx, down(vts, d->vts, 12 * (size_t)d->nv));
    VC_HIP(ctx, down(edges, v.E, 8 * (size_t)v.ne));
    VC_HIP(ctx, down(face_off, v.fo, 8 * ((size_t)d->nf + 1)));
    VC_HIP(ctx, down(face_vts, v.fv, 4 * (size_t)v.nfe));
    VC_HIP(ctx, down(face_sites, d->fs, 8 * (size_t)d->nf));
    VC_HIP(ctx, down(sites, d->sites, 12 * (size_t)d->ns));
    VC_HIP(ctx, down(radii, d->radii, 4 * (size_t)d->nv));
    VC_HIP(ctx, down(valid, d->valid, (size_t)d->nv));
    VC_HIP(ctx, hipStreamSynchronize(S));
    if (face_off && d->nf == 0)
        face_off[0] = 0;
    return VC_OK;
}

int vc_vdiag_extract_inside(vc_ctx* ctx, vc_vdiag* d, int all_valid,
                            int32_t* vts_idx, float* vm, int64_t vts_cap,
                            int64_t* n_vts, int32_t* edges, float* em,
                            int64_t edge_cap, int64_t* n_edges, int32_t* tris,
                            int32_t* from_fi, float* fm, int64_t tri_cap,
                            int64_t* n_tris, vc_cc** inside)
{
    if (check(ctx, d) != VC_OK || !n_vts || !n_edges || !n_tris)
        return VC_ERR_INVALID;
    if (inside)
        *inside = nullptr;
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    vc_cc_dview v;
    vc_cc_dview_get(d->cc, &v);
    if (3 * v.nfe >= (int64_t)INT32_MAX)
        return vc_fail(ctx, VC_ERR_INVALID, "vc_vdiag_extract_inside: complex too large");
    Tmp t;
    uint8_t *vval, *fval;
    float *lv, *le, *lf;
    int rc = VC_OK;
    auto get = [&](auto** p, size_t n) {
        if (rc == VC_OK)
            rc = t.get(ctx, p, n);
    };
    get(&vval, (size_t)d->nv + 1);
    get(&fval, (size_t)d->nf + 1);
    get(&lv, (size_t)d->nv + 1);
    get(&le, (size_t)v.ne + 1);
    get(&lf, (size_t)d->nf + 1);
    if (rc != VC_OK)
        return rc;
    if (d->nv)
        eff_valid<<<nblk(d->nv), kB, 0, S>>>(d->nv, d->valid, all_valid, vval);
    if (d->nf)
        face_valid<<<nblk(d->nf), kB, 0, S>>>(d->nf, v.fo, v.fv, vval, fval);
    VC_CHECK_LAUNCH(ctx);
    if ((rc = vc_voro_measures_dev(ctx, d->nv, v.ne, d->nf, d->sites, d->fs, fval,
                                   v.ve_off, v.ve, v.ef_off, v.ef, lv, le, lf)) != VC_OK)
        return rc;
    vc_voro_xi x;
    struct FreeXi
    {
        vc_voro_xi* x;
        ~FreeXi() { vc_voro_xi_free(x); }
    } fx{&x};
    if ((rc = vc_voro_extract_dev(ctx, d->nv, v.ne, d->nf, v.E, v.fo, v.fv, vval, le, lf,
                                  &x)) != VC_OK)
        return rc;
    *n_vts = x.nv;
    *n_edges = x.ne;
    *n_tris = x.nt;
    if ((vts_idx && vts_cap < x.nv) || (vm && vts_cap < x.nv) ||
        ((edges || em) && edge_cap < x.ne) || ((tris || from_fi || fm) && tri_cap < x.nt))
        return VC_ERR_CAPACITY;
    float* dvm = nullptr;
    if (vm && x.nv)
    {
        if ((rc = t.get(ctx, &dvm, (size_t)x.nv)) != VC_OK)
            return rc;
        gather_vm<<<nblk(x.nv), kB, 0, S>>>(x.nv, x.vi, lv, dvm);
        VC_CHECK_LAUNCH(ctx);
    }
    auto down = [&](void* h, const void* dp, size_t bytes) {
        return (h && bytes) ? hipMemcpyAsync(h, dp, bytes, hipMemcpyDeviceToHost, S)
                            : hipSuccess;
    };
    VC_HIP(ctx, down(vts_idx, x.vi, 4 * (size_t)x.nv));
    VC_HIP(ctx, down(vm, dvm, 4 * (size_t)x.nv));
    VC_HIP(ctx, down(edges, x.eo, 8 * (size_t)x.ne));
    VC_HIP(ctx, down(em, x.em, 4 * (size_t)x.ne));
    VC_HIP(ctx, down(tris, x.tr, 12 * (size_t)x.nt));
    VC_HIP(ctx, down(from_fi, x.ff, 4 * (size_t)x.nt));
    VC_HIP(ctx, down(fm, x.fm, 4 * (size_t)x.nt));
    if (inside)
    {
        int64_t* to = nullptr;
        if ((rc = t.get(ctx, &to, (size_t)x.nt + 1)) != VC_OK)
            return rc;
        offsets_step<<<nblk(x.nt + 1), kB, 0, S>>>(x.nt, 3, to);
        VC_CHECK_LAUNCH(ctx);
        if ((rc = vc_cc_create_dev(ctx, (int32_t)x.nv, (int32_t)x.ne, x.eo, (int32_t)x.nt, to,
                                   x.tr, 3 * x.nt, inside)) != VC_OK)
            return rc;
    }
    VC_HIP(ctx, hipStreamSynchronize(S));
    return VC_OK;
}

} // extern "C"
