// vc_batch.cpp -- batched volumes (C ABI of include/vc_batch.h) and the
// launch-parameter autotuner.
#include <algorithm>
#include <cstring>
#include <new>
#include <set>
#include <vector>

#include "vc_internal.h"

struct vc_batch
{
    vc_ctx* ctx = nullptr;
    vc_sched* s = nullptr;
};

namespace
{

std::vector<int64_t> shape_key(const vc_sched* s)
{
    int64_t amax = 0;
    for (const auto& d : s->h_desc)
        amax = std::max<int64_t>(amax, std::max(d.nx, std::max(d.ny, d.nz)));
    return {(int64_t)s->n, s->voxels, amax};
}

bool same(const vc_schedule& a, const vc_schedule& b)
{
    return std::memcmp(&a, &b, sizeof(a)) == 0;
}

int check_batch(vc_ctx* ctx, vc_batch* b)
{
    if (!ctx || !b || b->ctx != ctx)
        return VC_ERR_INVALID;
    return VC_OK;
}

int batch_thin(vc_ctx* ctx, vc_batch* b, const float* levels, int n_levels,
               vc_thin_stats* st)
{
    vc_sched* s = b->s;
    for (int i = 0; i < s->n; ++i)
    {
        int rc = vc_check_levels(ctx, levels + (size_t)i * n_levels, n_levels);
        if (rc != VC_OK)
            return vc_fail(ctx, rc, "volume %d: %s", i, ctx->err.c_str());
        if (s->vols[i]->measure_kind == 0)
            return vc_fail(ctx, VC_ERR_STATE,
                           "volume %d: run vc_batch_measure or "
                           "vc_set_measure_f32 first",
                           i);
    }
    return vc_thin_sched(ctx, s, levels, n_levels, st);
}

// GPU time of fn() on the context stream, median of reps after one warm-up
template <typename F>
int time_median(vc_ctx* ctx, int reps, F fn, float* out_ms)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    VC_HIP(ctx, hipEventCreate(&e0));
    hipError_t e = hipEventCreate(&e1);
    if (e != hipSuccess)
    {
        (void)hipEventDestroy(e0);
        return vc_fail(ctx, VC_ERR_HIP, "hipEventCreate: %s",
                       hipGetErrorString(e));
    }
    std::vector<float> t;
    int rc = fn(); // warm-up (captures graphs, first-touch)
    for (int r = 0; r < reps && rc == VC_OK; ++r)
    {
        e = hipEventRecord(e0, ctx->stream);
        if (e == hipSuccess && (rc = fn()) == VC_OK)
        {
            e = hipEventRecord(e1, ctx->stream);
            if (e == hipSuccess)
                e = hipEventSynchronize(e1);
            float ms = 0.0f;
            if (e == hipSuccess)
                e = hipEventElapsedTime(&ms, e0, e1);
            t.push_back(ms);
        }
        if (e != hipSuccess)
            rc = vc_fail(ctx, VC_ERR_HIP, "autotune timing: %s",
                         hipGetErrorString(e));
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != VC_OK)
        return rc;
    std::sort(t.begin(), t.end());
    *out_ms = t[t.size() / 2];
    return VC_OK;
}

} // namespace

extern "C" {

int vc_schedule_defaults(vc_schedule* out)
{
    if (!out)
        return VC_ERR_INVALID;
    *out = vc_schedule_default();
    return VC_OK;
}

int vc_batch_create(vc_ctx* ctx, vc_vol* const* vols, int n, vc_batch** out)
{
    if (!ctx || !vols || !out)
        return VC_ERR_INVALID;
    *out = nullptr;
    if (n < 1 || n > 65535)
        return vc_fail(ctx, VC_ERR_INVALID, "batch of %d volumes not in "
                                            "[1, 65535]", n);
    std::set<const vc_vol*> seen;
    for (int i = 0; i < n; ++i)
    {
        if (!vols[i] || vols[i]->ctx != ctx)
            return vc_fail(ctx, VC_ERR_INVALID,
                           "volume %d is null or of another context", i);
        if (!seen.insert(vols[i]).second)
            return vc_fail(ctx, VC_ERR_INVALID, "volume %d appears twice", i);
    }
    VC_HIP(ctx, hipSetDevice(ctx->device));
    vc_batch* b = new (std::nothrow) vc_batch();
    if (!b)
        return VC_ERR_NOMEM;
    b->ctx = ctx;
    int rc = vc_sched_build(ctx, vols, n, vc_schedule_default(), &b->s);
    if (rc == VC_OK)
    {
        auto it = ctx->tuned.find(shape_key(b->s));
        if (it != ctx->tuned.end())
            rc = vc_sched_apply(ctx, b->s, it->second);
    }
    if (rc != VC_OK)
    {
        vc_sched_free(b->s);
        delete b;
        return rc;
    }
    *out = b;
    return VC_OK;
}

int vc_batch_destroy(vc_ctx* ctx, vc_batch* b)
{
    int rc = check_batch(ctx, b);
    if (rc != VC_OK)
        return rc;
    (void)hipStreamSynchronize(ctx->stream);
    vc_sched_free(b->s);
    delete b;
    return VC_OK;
}

int vc_batch_size(const vc_batch* b, int* n, int64_t* voxels)
{
    if (!b || !n || !voxels)
        return VC_ERR_INVALID;
    *n = b->s->n;
    *voxels = b->s->voxels;
    return VC_OK;
}

int vc_batch_edt(vc_ctx* ctx, vc_batch* b, int flags)
{
    int rc = check_batch(ctx, b);
    if (rc != VC_OK)
        return rc;
    if (flags != 0)
        return vc_fail(ctx, VC_ERR_INVALID, "vc_batch_edt: flags must be 0");
    return vc_edt_sched(ctx, b->s);
}

int vc_batch_measure(vc_ctx* ctx, vc_batch* b)
{
    int rc = check_batch(ctx, b);
    if (rc != VC_OK)
        return rc;
    for (int i = 0; i < b->s->n; ++i)
        if (!b->s->vols[i]->edt_done)
            return vc_fail(ctx, VC_ERR_STATE,
                           "vc_batch_measure: volume %d has no EDT yet", i);
    return vc_measure_sched(ctx, b->s);
}

int vc_batch_thin(vc_ctx* ctx, vc_batch* b, const float* levels, int n_levels,
                  vc_thin_stats* stats)
{
    int rc = check_batch(ctx, b);
    if (rc != VC_OK)
        return rc;
    if (!levels)
        return VC_ERR_INVALID;
    return batch_thin(ctx, b, levels, n_levels, stats);
}

int vc_batch_run(vc_ctx* ctx, vc_batch* b, const float* levels, int n_levels,
                 vc_thin_stats* stats)
{
    int rc = check_batch(ctx, b);
    if (rc != VC_OK)
        return rc;
    if (!levels)
        return VC_ERR_INVALID;
    for (int i = 0; i < b->s->n; ++i)
        if ((rc = vc_check_levels(ctx, levels + (size_t)i * n_levels,
                                  n_levels)) != VC_OK)
            return rc;
    if ((rc = vc_edt_sched(ctx, b->s)) != VC_OK)
        return rc;
    if ((rc = vc_measure_sched(ctx, b->s)) != VC_OK)
        return rc;
    return batch_thin(ctx, b, levels, n_levels, stats);
}

int vc_batch_get_schedule(const vc_batch* b, vc_schedule* out)
{
    if (!b || !out)
        return VC_ERR_INVALID;
    *out = b->s->sc;
    return VC_OK;
}

int vc_batch_set_schedule(vc_ctx* ctx, vc_batch* b, const vc_schedule* sc)
{
    int rc = check_batch(ctx, b);
    if (rc != VC_OK)
        return rc;
    if (!sc)
        return VC_ERR_INVALID;
    return vc_sched_apply(ctx, b->s, *sc);
}

int vc_batch_tune(vc_ctx* ctx, vc_batch* b, const float* levels, int n_levels,
                  int reps, vc_schedule* chosen, float* times_ms)
{
    int rc = check_batch(ctx, b);
    if (rc != VC_OK)
        return rc;
    if (!levels || reps < 1 || reps > 100)
        return vc_fail(ctx, VC_ERR_INVALID, "vc_batch_tune: reps in [1, 100]");
    for (int i = 0; i < b->s->n; ++i)
        if ((rc = vc_check_levels(ctx, levels + (size_t)i * n_levels,
                                  n_levels)) != VC_OK)
            return rc;
    vc_sched* s = b->s;
    auto edt_measure = [&]() -> int {
        int r = vc_edt_sched(ctx, s);
        return r == VC_OK ? vc_measure_sched(ctx, s) : r;
    };
    auto thin = [&]() -> int {
        return vc_thin_sched(ctx, s, levels, n_levels, nullptr);
    };
    const std::vector<int64_t> key = shape_key(s);
    float t_front = 0.0f, t_thin = 0.0f;
    auto cached = ctx->tuned.find(key);
    if (cached != ctx->tuned.end())
    {
        if ((rc = vc_sched_apply(ctx, s, cached->second)) != VC_OK)
            return rc;
        if ((rc = time_median(ctx, 1, edt_measure, &t_front)) != VC_OK)
            return rc;
        if ((rc = time_median(ctx, 1, thin, &t_thin)) != VC_OK)
            return rc;
    }
    else
    {
        // EDT + measure candidates, incumbent first: (lds KiB, grid per CU)
        static const int kEdt[][2] = {{16, 0}, {32, 0}, {8, 0},
                                      {64, 0}, {16, 8}, {32, 8}};
        vc_schedule best = vc_schedule_default();
        float best_t = 0.0f;
        for (size_t c = 0; c < sizeof(kEdt) / sizeof(kEdt[0]); ++c)
        {
            vc_schedule sc = best;
            sc.edt_lds_kb = kEdt[c][0];
            sc.edt_grid_per_cu = kEdt[c][1];
            if ((rc = vc_sched_apply(ctx, s, sc)) != VC_OK)
                return rc;
            float t = 0.0f;
            if ((rc = time_median(ctx, reps, edt_measure, &t)) != VC_OK)
                return rc;
            if (c == 0 || t < 0.98f * best_t)
            {
                best = sc;
                best_t = t;
            }
        }
        t_front = best_t;
        // thinning candidates: (grid per CU, graph, sweeps per sync, first
        // tile sweep); tile sweeps hold 37 KB of LDS, four workgroups per CU
        static const int kThin[][4] = {{0, 0, 0, 0},  {0, 0, -1, 0},
                                       {0, 1, 0, 0},
                                       {0, 0, 16, 0}, {2, 0, 0, 0},
                                       {0, 0, 0, 4},  {0, 0, 0, -1},
                                       {0, 1, 8, -1}, {8, 0, 0, -1},
                                       {8, 1, 16, -1}};
        vc_schedule incumbent = best;
        best_t = 0.0f;
        for (size_t c = 0; c < sizeof(kThin) / sizeof(kThin[0]); ++c)
        {
            vc_schedule sc = incumbent;
            sc.thin_grid_per_cu = kThin[c][0];
            sc.use_graph = kThin[c][1];
            sc.sweeps_per_sync = kThin[c][2];
            sc.tile_sweep = kThin[c][3];
            if ((rc = vc_sched_apply(ctx, s, sc)) != VC_OK)
                return rc;
            float t = 0.0f;
            if ((rc = time_median(ctx, reps, thin, &t)) != VC_OK)
                return rc;
            if (c == 0 || t < 0.98f * best_t)
            {
                best = sc;
                best_t = t;
            }
        }
        t_thin = best_t;
        if (!same(best, s->sc) && (rc = vc_sched_apply(ctx, s, best)) != VC_OK)
            return rc;
        // leave the volumes thinned under the installed schedule
        if ((rc = thin()) != VC_OK)
            return rc;
        ctx->tuned[key] = best;
    }
    VC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (chosen)
        *chosen = s->sc;
    if (times_ms)
    {
        times_ms[0] = t_front;
        times_ms[1] = t_thin;
    }
    return VC_OK;
}

} // extern "C"
