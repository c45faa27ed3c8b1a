// vc_core.hip -- context, volumes, results and error handling of the C ABI.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>

#include "vc_internal.h"

int vc_fail(vc_ctx* ctx, int code, const char* fmt, ...)
{
    if (ctx)
    {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

int vc_alloc(vc_ctx* ctx, void** p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0)
        bytes = 256;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess)
    {
        *p = nullptr;
        (void)hipGetLastError();
        return vc_fail(ctx, VC_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s",
                       bytes, hipGetErrorString(e));
    }
    return VC_OK;
}

void vc_free(void* p)
{
    if (p)
        (void)hipFree(p);
}

int vc_ws_get(vc_ctx* ctx, int slot, size_t bytes, void** p)
{
    *p = nullptr;
    if (slot < 0 || slot >= VC_WS_SLOTS)
        return vc_fail(ctx, VC_ERR_INVALID, "scratch slot %d", slot);
    if (ctx->ws_bytes[slot] < bytes || !ctx->ws[slot])
    {
        // hipFree waits for the device, so the buffer's last user is done
        vc_free(ctx->ws[slot]);
        ctx->ws[slot] = nullptr;
        ctx->ws_bytes[slot] = 0;
        // a little headroom, so slowly growing inputs do not reallocate
        // every call
        const size_t want = bytes + bytes / 8;
        const int rc = vc_alloc(ctx, &ctx->ws[slot], want);
        if (rc != VC_OK)
            return rc;
        ctx->ws_bytes[slot] = want;
    }
    *p = ctx->ws[slot];
    return VC_OK;
}

int64_t vc_level_threshold(float level)
{
    double l = (double)level;
    if (!(l > 0.0))
        return 0;
    double l2 = l * l; // exact: a float squared fits in a double mantissa
    if (l2 > 9.0e18)
        return INT64_MAX;
    return (int64_t)std::ceil(l2);
}

int vc_check_levels(vc_ctx* ctx, const float* levels, int n)
{
    if (n < 1 || n > VC_MAX_LEVELS)
        return vc_fail(ctx, VC_ERR_INVALID, "n_levels %d not in [1, %d]", n,
                       VC_MAX_LEVELS);
    for (int k = 0; k < n; ++k)
    {
        if (!(levels[k] == levels[k]))
            return vc_fail(ctx, VC_ERR_INVALID, "level %d is NaN", k);
        if (k > 0 && levels[k] < levels[k - 1])
            return vc_fail(ctx, VC_ERR_INVALID,
                           "levels must be ascending (level %d < level %d)", k,
                           k - 1);
    }
    return VC_OK;
}

static int bits_for(int n)
{
    int b = 0;
    while ((1 << b) < n)
        b++;
    return b;
}

extern "C" {

int vc_ctx_create(int device, vc_ctx** out)
{
    if (!out)
        return VC_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    {
        (void)hipGetLastError();
        return VC_ERR_HIP;
    }
    if (device < 0 || device >= ndev)
        return VC_ERR_INVALID;
    vc_ctx* ctx = new (std::nothrow) vc_ctx();
    if (!ctx)
        return VC_ERR_NOMEM;
    ctx->device = device;
    int rc = VC_OK;
    if (hipSetDevice(device) != hipSuccess)
        rc = VC_ERR_HIP;
    if (rc == VC_OK)
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) != hipSuccess)
            rc = VC_ERR_HIP;
        else
            ctx->cu_count = prop.multiProcessorCount;
    }
    if (rc == VC_OK && hipStreamCreateWithFlags(&ctx->stream,
                                                hipStreamNonBlocking) !=
                           hipSuccess)
        rc = VC_ERR_HIP;
    if (rc == VC_OK &&
        (hipStreamCreateWithFlags(&ctx->stream_aux, hipStreamNonBlocking) != hipSuccess ||
         hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
         hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess))
        rc = VC_ERR_HIP;
    if (rc == VC_OK)
        rc = vc_lut_build(ctx);
    if (rc == VC_OK)
        rc = vc_noalloc_setup(ctx);
    if (rc == VC_OK && hipStreamSynchronize(ctx->stream) != hipSuccess)
        rc = VC_ERR_HIP;
    if (rc != VC_OK)
    {
        vc_ctx_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return VC_OK;
}

int vc_ctx_destroy(vc_ctx* ctx)
{
    if (!ctx)
        return VC_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream)
        (void)hipStreamSynchronize(ctx->stream);
    vc_free(ctx->d_lut);
    for (int k = 0; k < VC_WS_SLOTS; ++k)
        vc_free(ctx->ws[k]);
    vc_noalloc_teardown(ctx);
    for (hipEvent_t e : ctx->ev_free)
        (void)hipEventDestroy(e);
    for (auto& pr : ctx->ev_open)
    {
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
    }
    if (ctx->stream_aux)
    {
        (void)hipStreamSynchronize(ctx->stream_aux);
        (void)hipStreamDestroy(ctx->stream_aux);
    }
    if (ctx->ev_fork)
        (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join)
        (void)hipEventDestroy(ctx->ev_join);
    if (ctx->stream)
        (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return VC_OK;
}

int vc_last_error(const vc_ctx* ctx, char* buf, size_t n)
{
    if (!ctx || !buf || n == 0)
        return VC_ERR_INVALID;
    std::snprintf(buf, n, "%s", ctx->err.c_str());
    return VC_OK;
}

int vc_ctx_stream(vc_ctx* ctx, void** s)
{
    if (!ctx || !s)
        return VC_ERR_INVALID;
    *s = (void*)ctx->stream;
    return VC_OK;
}

int vc_ctx_synchronize(vc_ctx* ctx)
{
    if (!ctx)
        return VC_ERR_INVALID;
    VC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VC_OK;
}

int vc_simple_table_host(vc_ctx* ctx, uint32_t* words)
{
    if (!ctx || !words)
        return VC_ERR_INVALID;
    VC_HIP(ctx, hipMemcpyAsync(words, ctx->d_lut, sizeof(uint32_t) << 21,
                               hipMemcpyDeviceToHost, ctx->stream));
    VC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VC_OK;
}

static int volume_new(vc_ctx* ctx, int nx, int ny, int nz, vc_vol** out)
{
    if (!ctx || !out)
        return VC_ERR_INVALID;
    *out = nullptr;
    if (nx <= 0 || ny <= 0 || nz <= 0 || nx > VC_MAX_AXIS ||
        ny > VC_MAX_AXIS || nz > VC_MAX_AXIS)
        return vc_fail(ctx, VC_ERR_INVALID,
                       "volume dims %d x %d x %d out of range [1, %d]", nx, ny,
                       nz, VC_MAX_AXIS);
    int64_t N = (int64_t)nx * ny * nz;
    int bx = bits_for(nx), by = bits_for(ny), bz = bits_for(nz);
    if (N >= (int64_t(1) << 31) || bx + by + bz > 31)
        return vc_fail(ctx, VC_ERR_INVALID,
                       "volume %d x %d x %d too large (N < 2^31 and packed "
                       "coordinates <= 31 bits required)",
                       nx, ny, nz);
    double maxd2 = (double)(nx - 1) * (nx - 1) + (double)(ny - 1) * (ny - 1) +
                   (double)(nz - 1) * (nz - 1);
    if (maxd2 > 2147483647.0)
        return vc_fail(ctx, VC_ERR_INVALID,
                       "volume %d x %d x %d: squared diagonal exceeds int32",
                       nx, ny, nz);
    vc_vol* v = new (std::nothrow) vc_vol();
    if (!v)
        return VC_ERR_NOMEM;
    v->ctx = ctx;
    v->nx = nx;
    v->ny = ny;
    v->nz = nz;
    v->N = N;
    v->pk.bx = bx;
    v->pk.by = by;
    v->pk.bz = bz;
    v->pk.my = (1u << by) - 1u;
    v->pk.mz = (1u << bz) - 1u;
    int rc = vc_alloc(ctx, (void**)&v->occ, (size_t)N);
    if (rc != VC_OK)
    {
        delete v;
        return rc;
    }
    *out = v;
    return VC_OK;
}

int vc_volume_upload(vc_ctx* ctx, const uint8_t* occ, int nx, int ny, int nz,
                     vc_vol** out)
{
    if (!ctx)
        return VC_ERR_INVALID;
    if (!occ)
        return vc_fail(ctx, VC_ERR_EMPTY_VOL, "null volume");
    VC_HIP(ctx, hipSetDevice(ctx->device));
    int rc = volume_new(ctx, nx, ny, nz, out);
    if (rc != VC_OK)
        return rc;
    hipError_t e = hipMemcpyAsync((*out)->occ, occ, (size_t)(*out)->N,
                                  hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess)
    {
        vc_volume_release(ctx, *out);
        *out = nullptr;
        return vc_fail(ctx, VC_ERR_HIP, "volume upload: %s",
                       hipGetErrorString(e));
    }
    return VC_OK;
}

int vc_volume_from_device(vc_ctx* ctx, const uint8_t* d_occ, int nx, int ny,
                          int nz, vc_vol** out)
{
    if (!ctx)
        return VC_ERR_INVALID;
    if (!d_occ)
        return vc_fail(ctx, VC_ERR_EMPTY_VOL, "null volume");
    VC_HIP(ctx, hipSetDevice(ctx->device));
    int rc = volume_new(ctx, nx, ny, nz, out);
    if (rc != VC_OK)
        return rc;
    hipError_t e = hipMemcpyAsync((*out)->occ, d_occ, (size_t)(*out)->N,
                                  hipMemcpyDeviceToDevice, ctx->stream);
    if (e != hipSuccess)
    {
        vc_volume_release(ctx, *out);
        *out = nullptr;
        return vc_fail(ctx, VC_ERR_HIP, "volume copy: %s",
                       hipGetErrorString(e));
    }
    return VC_OK;
}

int vc_volume_release(vc_ctx* ctx, vc_vol* v)
{
    if (!ctx || !v)
        return VC_ERR_INVALID;
    (void)hipStreamSynchronize(ctx->stream);
    vc_free(v->occ);
    vc_free(v->edt);
    vc_free(v->ft);
    vc_free(v->lam2);
    vc_free(v->meas_f);
    vc_free(v->lvl);
    vc_free(v->label);
    vc_free(v->alive);
    vc_free(v->elig);
    vc_free(v->dflag);
    vc_free(v->alive2);
    vc_free(v->tchg);
    vc_free(v->eflag);
    vc_sched_free(v->self);
    vc_free(v->bnd_sites);
    vc_free(v->bnd_quads);
    delete v;
    return VC_OK;
}

int vc_volume_dims(const vc_vol* v, int* nx, int* ny, int* nz)
{
    if (!v || !nx || !ny || !nz)
        return VC_ERR_INVALID;
    *nx = v->nx;
    *ny = v->ny;
    *nz = v->nz;
    return VC_OK;
}

int vc_edt(vc_ctx* ctx, vc_vol* vol, int flags)
{
    if (!ctx || !vol || vol->ctx != ctx)
        return VC_ERR_INVALID;
    if (flags != 0)
        return vc_fail(ctx, VC_ERR_INVALID, "vc_edt: flags must be 0");
    int rc = vc_vol_sched(ctx, vol);
    return rc != VC_OK ? rc : vc_edt_sched(ctx, vol->self);
}

int vc_measure(vc_ctx* ctx, vc_vol* vol)
{
    if (!ctx || !vol || vol->ctx != ctx)
        return VC_ERR_INVALID;
    if (!vol->edt_done)
        return vc_fail(ctx, VC_ERR_STATE, "vc_measure: run vc_edt first");
    int rc = vc_vol_sched(ctx, vol); // the EDT may have come from a batch
    return rc != VC_OK ? rc : vc_measure_sched(ctx, vol->self);
}

int vc_set_measure_f32(vc_ctx* ctx, vc_vol* vol, const float* d_measure)
{
    if (!ctx || !vol || vol->ctx != ctx || !d_measure)
        return VC_ERR_INVALID;
    if (!vol->meas_f)
    {
        int rc = vc_alloc(ctx, (void**)&vol->meas_f,
                          sizeof(float) * (size_t)vol->N);
        if (rc != VC_OK)
            return rc;
    }
    VC_HIP(ctx, hipMemcpyAsync(vol->meas_f, d_measure,
                               sizeof(float) * (size_t)vol->N,
                               hipMemcpyDeviceToDevice, ctx->stream));
    vol->measure_kind = 2;
    return VC_OK;
}

int vc_thin_levels(vc_ctx* ctx, vc_vol* vol, const float* levels, int n,
                   vc_thin_stats* st)
{
    if (!ctx || !vol || vol->ctx != ctx || !levels)
        return VC_ERR_INVALID;
    int rc = vc_check_levels(ctx, levels, n);
    if (rc != VC_OK)
        return rc;
    if (vol->measure_kind == 0)
        return vc_fail(ctx, VC_ERR_STATE,
                       "vc_thin_levels: run vc_measure or vc_set_measure_f32");
    rc = vc_vol_sched(ctx, vol);
    return rc != VC_OK ? rc : vc_thin_sched(ctx, vol->self, levels, n, st);
}

int vc_boundary_extract(vc_ctx* ctx, vc_vol* vol, float* sites,
                        int64_t site_cap, int64_t* n_sites, int32_t* quads,
                        int64_t quad_cap, int64_t* n_quads, vc_euler* euler)
{
    if (!ctx)
        return VC_ERR_INVALID;
    if (!vol)
        return vc_fail(ctx, VC_ERR_EMPTY_VOL, "null volume");
    if (vol->ctx != ctx || !n_sites || !n_quads)
        return VC_ERR_INVALID;
    return vc_boundary_run(ctx, vol, sites, site_cap, n_sites, quads, quad_cap,
                           n_quads, euler);
}

static int buf_info(vc_ctx* ctx, vc_vol* v, int which, void** p, size_t* esz)
{
    switch (which)
    {
    case VC_BUF_OCC: *p = v->occ; *esz = 1; break;
    case VC_BUF_EDT: *p = v->edt; *esz = 4; break;
    case VC_BUF_FT: *p = v->ft; *esz = 4; break;
    case VC_BUF_LAM2: *p = v->lam2; *esz = 4; break;
    case VC_BUF_LABEL: *p = v->label; *esz = 1; break;
    case VC_BUF_LEVEL: *p = v->lvl; *esz = 1; break;
    default: return vc_fail(ctx, VC_ERR_INVALID, "unknown buffer %d", which);
    }
    if (!*p)
        return vc_fail(ctx, VC_ERR_STATE, "buffer %d not computed yet", which);
    return VC_OK;
}

int vc_get_device(vc_ctx* ctx, vc_vol* v, int which, void* dst, int64_t bytes)
{
    if (!ctx || !v || v->ctx != ctx || !dst)
        return VC_ERR_INVALID;
    void* src;
    size_t esz;
    int rc = buf_info(ctx, v, which, &src, &esz);
    if (rc != VC_OK)
        return rc;
    if (bytes != (int64_t)(esz * (size_t)v->N))
        return vc_fail(ctx, VC_ERR_INVALID, "vc_get: %lld bytes, need %lld",
                       (long long)bytes, (long long)(esz * (size_t)v->N));
    if (which == VC_BUF_FT)
        return vc_ft_to_linear(ctx, v, (int32_t*)dst);
    VC_HIP(ctx, hipMemcpyAsync(dst, src, (size_t)bytes,
                               hipMemcpyDeviceToDevice, ctx->stream));
    return VC_OK;
}

int vc_get(vc_ctx* ctx, vc_vol* v, int which, void* dst, int64_t bytes)
{
    if (!ctx || !v || v->ctx != ctx || !dst)
        return VC_ERR_INVALID;
    void* src;
    size_t esz;
    int rc = buf_info(ctx, v, which, &src, &esz);
    if (rc != VC_OK)
        return rc;
    if (bytes != (int64_t)(esz * (size_t)v->N))
        return vc_fail(ctx, VC_ERR_INVALID, "vc_get: %lld bytes, need %lld",
                       (long long)bytes, (long long)(esz * (size_t)v->N));
    if (which == VC_BUF_FT)
    {
        int32_t* tmp = nullptr;
        rc = vc_alloc(ctx, (void**)&tmp, (size_t)bytes);
        if (rc != VC_OK)
            return rc;
        rc = vc_ft_to_linear(ctx, v, tmp);
        if (rc == VC_OK)
        {
            hipError_t e = hipMemcpyAsync(dst, tmp, (size_t)bytes,
                                          hipMemcpyDeviceToHost, ctx->stream);
            if (e == hipSuccess)
                e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess)
                rc = vc_fail(ctx, VC_ERR_HIP, "vc_get(FT): %s",
                             hipGetErrorString(e));
        }
        vc_free(tmp);
        return rc;
    }
    VC_HIP(ctx, hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost,
                               ctx->stream));
    VC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VC_OK;
}

} // extern "C"
