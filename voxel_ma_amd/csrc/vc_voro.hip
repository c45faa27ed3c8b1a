// vc_voro.hip -- inside tagging of points and lambda measures of a 2-complex
// on the device (include/vc_voro.h), gfx950.
//
// Both are gathers: one thread per point / face / edge / vertex, the measure
// of an edge or vertex a maximum over its CSR neighbourhood.
#include <cmath>
#include <vector>

#include "vc_internal.h"
#include "../../include/vc_voro.h"

namespace
{

constexpr int kBlock = 256;

// std::round (half away from zero) then the occupancy of that voxel; points
// whose rounded coordinate does not fit an int are outside
__global__ void __launch_bounds__(kBlock)
    tag_points(const float* __restrict__ pts, int64_t n,
               const uint8_t* __restrict__ occ, int nx, int ny, int nz,
               uint8_t* __restrict__ valid)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int v[3];
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        const float r = roundf(pts[3 * i + a]);
        ok = ok && (r > -1.0e9f && r < 1.0e9f); // also false for NaN
        v[a] = ok ? (int)r : 0;
    }
    ok = ok && (unsigned)v[0] < (unsigned)nx && (unsigned)v[1] < (unsigned)ny &&
         (unsigned)v[2] < (unsigned)nz;
    valid[i] = ok && occ[((int64_t)v[0] * ny + v[1]) * nz + v[2]] != 0;
}

// |a - b| as trimesh::dist: ((d0^2 + d1^2) + d2^2), no contraction,
// correctly rounded square root
__device__ float site_dist(const float* __restrict__ s, int a, int b)
{
#pragma clang fp contract(off)
    const float d0 = s[3 * b] - s[3 * a];
    const float d1 = s[3 * b + 1] - s[3 * a + 1];
    const float d2 = s[3 * b + 2] - s[3 * a + 2];
    const float q = (d0 * d0 + d1 * d1) + d2 * d2;
    return sqrtf(q);
}

__global__ void __launch_bounds__(kBlock)
    face_lambda(const float* __restrict__ sites,
                const int32_t* __restrict__ face_sites, int32_t nf,
                float* __restrict__ lam_f)
{
    const int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf)
        return;
    lam_f[f] = site_dist(sites, face_sites[2 * f], face_sites[2 * f + 1]);
}

__global__ void __launch_bounds__(kBlock)
    edge_lambda(const float* __restrict__ lam_f,
                const uint8_t* __restrict__ face_valid,
                const int64_t* __restrict__ ef_off,
                const int32_t* __restrict__ ef, int32_t ne,
                float* __restrict__ lam_e)
{
    const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne)
        return;
    float m = 0.0f;
    for (int64_t j = ef_off[e]; j < ef_off[e + 1]; ++j)
    {
        const int32_t f = ef[j];
        if (face_valid[f])
            m = fmaxf(m, lam_f[f]);
    }
    lam_e[e] = m;
}

__global__ void __launch_bounds__(kBlock)
    vertex_lambda(const float* __restrict__ lam_f,
                  const uint8_t* __restrict__ face_valid,
                  const int64_t* __restrict__ ve_off,
                  const int32_t* __restrict__ ve,
                  const int64_t* __restrict__ ef_off,
                  const int32_t* __restrict__ ef, int32_t nv,
                  float* __restrict__ lam_v)
{
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv)
        return;
    float m = 0.0f;
    for (int64_t i = ve_off[v]; i < ve_off[v + 1]; ++i)
    {
        const int32_t e = ve[i];
        for (int64_t j = ef_off[e]; j < ef_off[e + 1]; ++j)
        {
            const int32_t f = ef[j];
            if (face_valid[f])
                m = fmaxf(m, lam_f[f]);
        }
    }
    lam_v[v] = m;
}

// last[v] = the highest face id whose polygon holds v (last starts at -1)
__global__ void __launch_bounds__(kBlock)
    face_last(const int64_t* __restrict__ face_off,
              const int32_t* __restrict__ face_vts, int32_t nf,
              int32_t* __restrict__ last)
{
    const int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf)
        return;
    for (int64_t j = face_off[f]; j < face_off[f + 1]; ++j)
        atomicMax(&last[face_vts[j]], f);
}

// r[v] = |v - first site of face last[v]| (trimesh::dist order), 0 if v is
// in no face
__global__ void __launch_bounds__(kBlock)
    vertex_radius(const float* __restrict__ vts, const float* __restrict__ sites,
                  const int32_t* __restrict__ face_sites,
                  const int32_t* __restrict__ last, int32_t nv,
                  float* __restrict__ r)
{
#pragma clang fp contract(off)
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv)
        return;
    const int32_t f = last[v];
    if (f < 0)
    {
        r[v] = 0.0f;
        return;
    }
    const int32_t s = face_sites[2 * f];
    const float d0 = vts[3 * v] - sites[3 * s];
    const float d1 = vts[3 * v + 1] - sites[3 * s + 1];
    const float d2 = vts[3 * v + 2] - sites[3 * s + 2];
    r[v] = sqrtf((d0 * d0 + d1 * d1) + d2 * d2);
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// device copies that are freed on every exit path
struct Scratch
{
    std::vector<void*> p;
    ~Scratch()
    {
        for (void* q : p)
            vc_free(q);
    }
    template <typename T>
    int get(vc_ctx* ctx, T** out, size_t count)
    {
        void* q = nullptr;
        int rc = vc_alloc(ctx, &q, count * sizeof(T) + 16);
        if (rc == VC_OK)
            p.push_back(q);
        *out = static_cast<T*>(q);
        return rc;
    }
};

} // namespace

extern "C" {

int vc_voro_tag_points(vc_ctx* ctx, vc_vol* vol, const float* pts, int64_t n,
                       uint8_t* valid)
{
    if (!ctx || !vol || vol->ctx != ctx || n < 0 || (n > 0 && (!pts || !valid)))
        return VC_ERR_INVALID;
    if (n == 0)
        return VC_OK;
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    Scratch sc;
    float* dp;
    uint8_t* dv;
    int rc = sc.get(ctx, &dp, 3 * (size_t)n);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dv, (size_t)n);
    if (rc != VC_OK)
        return rc;
    VC_HIP(ctx, hipMemcpyAsync(dp, pts, 12 * (size_t)n, hipMemcpyHostToDevice,
                               S));
    tag_points<<<blocks_for(n), kBlock, 0, S>>>(dp, n, vol->occ, vol->nx,
                                                 vol->ny, vol->nz, dv);
    VC_CHECK_LAUNCH(ctx);
    VC_HIP(ctx, hipMemcpyAsync(valid, dv, (size_t)n, hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    return VC_OK;
}

int vc_voro_measures(vc_ctx* ctx, int32_t nv, int32_t ne, int32_t nf,
                     const float* sites, int32_t nsites,
                     const int32_t* face_sites, const uint8_t* face_valid,
                     const int64_t* ve_off, const int32_t* ve,
                     const int64_t* ef_off, const int32_t* ef, float* lam_v,
                     float* lam_e, float* lam_f)
{
    if (!ctx || nv < 0 || ne < 0 || nf < 0 || nsites < 0 ||
        (nf > 0 && (!face_sites || !face_valid || !sites)) ||
        (nv > 0 && (!ve_off || (ve_off[nv] > 0 && !ve))) ||
        (ne > 0 && (!ef_off || (ef_off[ne] > 0 && !ef))))
        return VC_ERR_INVALID;
    for (int64_t f = 0; f < 2LL * nf; ++f)
        if (face_sites[f] < 0 || face_sites[f] >= nsites)
            return vc_fail(ctx, VC_ERR_INVALID, "face site %d out of range",
                           face_sites[f]);
    const int64_t nve = nv > 0 ? ve_off[nv] : 0, nef = ne > 0 ? ef_off[ne] : 0;
    for (int64_t i = 0; i < nve; ++i)
        if (ve[i] < 0 || ve[i] >= ne)
            return vc_fail(ctx, VC_ERR_INVALID, "vertex-edge entry out of range");
    for (int64_t i = 0; i < nef; ++i)
        if (ef[i] < 0 || ef[i] >= nf)
            return vc_fail(ctx, VC_ERR_INVALID, "edge-face entry out of range");
    for (int32_t v = 0; v < nv; ++v)
        if (ve_off[v] > ve_off[v + 1] || ve_off[v] < 0)
            return vc_fail(ctx, VC_ERR_INVALID, "vertex offsets not ascending");
    for (int32_t e = 0; e < ne; ++e)
        if (ef_off[e] > ef_off[e + 1] || ef_off[e] < 0)
            return vc_fail(ctx, VC_ERR_INVALID, "edge offsets not ascending");
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    Scratch sc;
    float *ds, *dlf, *dle, *dlv;
    int32_t *dfs, *dve, *def;
    uint8_t* dfv;
    int64_t *dveo, *defo;
    int rc = VC_OK;
    if (rc == VC_OK)
        rc = sc.get(ctx, &ds, 3 * (size_t)nsites);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dlf, (size_t)nf);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dle, (size_t)ne);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dlv, (size_t)nv);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dfs, 2 * (size_t)nf);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dfv, (size_t)nf);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dveo, (size_t)nv + 1);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dve, (size_t)nve);
    if (rc == VC_OK)
        rc = sc.get(ctx, &defo, (size_t)ne + 1);
    if (rc == VC_OK)
        rc = sc.get(ctx, &def, (size_t)nef);
    if (rc != VC_OK)
        return rc;
    auto up = [&](void* d, const void* h, size_t bytes) {
        return bytes ? hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, S)
                     : hipSuccess;
    };
    VC_HIP(ctx, up(ds, sites, 12 * (size_t)nsites));
    VC_HIP(ctx, up(dfs, face_sites, 8 * (size_t)nf));
    VC_HIP(ctx, up(dfv, face_valid, (size_t)nf));
    if (nv > 0)
    {
        VC_HIP(ctx, up(dveo, ve_off, 8 * ((size_t)nv + 1)));
        VC_HIP(ctx, up(dve, ve, 4 * (size_t)nve));
    }
    if (ne > 0)
    {
        VC_HIP(ctx, up(defo, ef_off, 8 * ((size_t)ne + 1)));
        VC_HIP(ctx, up(def, ef, 4 * (size_t)nef));
    }
    if (nf > 0)
    {
        face_lambda<<<blocks_for(nf), kBlock, 0, S>>>(ds, dfs, nf, dlf);
        VC_CHECK_LAUNCH(ctx);
    }
    if (ne > 0)
    {
        edge_lambda<<<blocks_for(ne), kBlock, 0, S>>>(dlf, dfv, defo, def, ne,
                                                       dle);
        VC_CHECK_LAUNCH(ctx);
    }
    if (nv > 0 && ne > 0)
    {
        vertex_lambda<<<blocks_for(nv), kBlock, 0, S>>>(dlf, dfv, dveo, dve,
                                                         defo, def, nv, dlv);
        VC_CHECK_LAUNCH(ctx);
    }
    else if (nv > 0)
        VC_HIP(ctx, hipMemsetAsync(dlv, 0, 4 * (size_t)nv, S));
    if (lam_f && nf > 0)
        VC_HIP(ctx, hipMemcpyAsync(lam_f, dlf, 4 * (size_t)nf,
                                   hipMemcpyDeviceToHost, S));
    if (lam_e && ne > 0)
        VC_HIP(ctx, hipMemcpyAsync(lam_e, dle, 4 * (size_t)ne,
                                   hipMemcpyDeviceToHost, S));
    if (lam_v && nv > 0)
        VC_HIP(ctx, hipMemcpyAsync(lam_v, dlv, 4 * (size_t)nv,
                                   hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    return VC_OK;
}

int vc_voro_radii(vc_ctx* ctx, int32_t nv, int32_t nf, const float* vts,
                  const int64_t* face_off, const int32_t* face_vts,
                  const float* sites, int32_t nsites, const int32_t* face_sites,
                  float* radii)
{
    if (!ctx || nv < 0 || nf < 0 || nsites < 0 || (nv > 0 && (!vts || !radii)) ||
        (nf > 0 && (!face_off || !face_vts || !face_sites || !sites)))
        return VC_ERR_INVALID;
    if (nf > 0 && face_off[0] != 0)
        return vc_fail(ctx, VC_ERR_INVALID, "face offsets must start at 0");
    for (int32_t f = 0; f < nf; ++f)
    {
        if (face_off[f + 1] < face_off[f])
            return vc_fail(ctx, VC_ERR_INVALID, "face offsets not ascending");
        if (face_sites[2 * f] < 0 || face_sites[2 * f] >= nsites ||
            face_sites[2 * f + 1] < 0 || face_sites[2 * f + 1] >= nsites)
            return vc_fail(ctx, VC_ERR_INVALID, "face site out of range");
    }
    const int64_t nfv = nf > 0 ? face_off[nf] : 0;
    for (int64_t j = 0; j < nfv; ++j)
        if (face_vts[j] < 0 || face_vts[j] >= nv)
            return vc_fail(ctx, VC_ERR_INVALID, "face vertex out of range");
    if (nv == 0)
        return VC_OK;
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    Scratch sc;
    float *dv, *ds, *dr;
    int32_t *dfv, *dfs, *dlast;
    int64_t* dfo;
    int rc = sc.get(ctx, &dv, 3 * (size_t)nv);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dr, (size_t)nv);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dlast, (size_t)nv);
    if (rc == VC_OK)
        rc = sc.get(ctx, &ds, 3 * (size_t)nsites);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dfs, 2 * (size_t)nf);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dfo, (size_t)nf + 1);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dfv, (size_t)nfv);
    if (rc != VC_OK)
        return rc;
    VC_HIP(ctx, hipMemcpyAsync(dv, vts, 12 * (size_t)nv, hipMemcpyHostToDevice, S));
    VC_HIP(ctx, hipMemsetAsync(dlast, 0xFF, 4 * (size_t)nv, S));
    if (nf > 0)
    {
        if (nsites > 0)
            VC_HIP(ctx, hipMemcpyAsync(ds, sites, 12 * (size_t)nsites,
                                       hipMemcpyHostToDevice, S));
        VC_HIP(ctx, hipMemcpyAsync(dfs, face_sites, 8 * (size_t)nf,
                                   hipMemcpyHostToDevice, S));
        VC_HIP(ctx, hipMemcpyAsync(dfo, face_off, 8 * ((size_t)nf + 1),
                                   hipMemcpyHostToDevice, S));
        if (nfv > 0)
            VC_HIP(ctx, hipMemcpyAsync(dfv, face_vts, 4 * (size_t)nfv,
                                       hipMemcpyHostToDevice, S));
        face_last<<<blocks_for(nf), kBlock, 0, S>>>(dfo, dfv, nf, dlast);
        VC_CHECK_LAUNCH(ctx);
    }
    vertex_radius<<<blocks_for(nv), kBlock, 0, S>>>(dv, ds, dfs, dlast, nv, dr);
    VC_CHECK_LAUNCH(ctx);
    VC_HIP(ctx, hipMemcpyAsync(radii, dr, 4 * (size_t)nv, hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    return VC_OK;
}

} // extern "C"
