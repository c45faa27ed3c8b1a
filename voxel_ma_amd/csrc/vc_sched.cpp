// vc_sched.cpp -- work tables of the volume kernels (host side).
//
// A launch of a volume kernel processes `total` work items; item b is one
// tile of one volume.  For several volumes the table d_items[kind][b] holds
// (volume << 32 | local tile); the tiles of a volume are consecutive, volumes
// in batch order.  For one volume there is no table (item = local tile).
#include <algorithm>
#include <cstring>
#include <new>

#include "vc_internal.h"

namespace
{

constexpr int kRowsZ = 4;       // EDT pass z: rows (wavefronts) per item
constexpr int kLiWaves = 4;     // level set-up: wavefronts per item
constexpr int kTpWords = 256;   // subfield pass: words per item
constexpr int kTileXY = 32;     // tile sweep: rows per tile along x and y

int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// z columns per tile: a power of two <= 64 whose tile fits the LDS budget
// (at least 1) and that does not exceed the row length by more than 2x
int pick_tile(int n_col, size_t elem, size_t budget, int nz)
{
    int tz = 64;
    while (tz > 1 && (size_t)tz * (size_t)n_col * elem > budget)
        tz >>= 1;
    while (tz > 1 && tz / 2 >= nz)
        tz >>= 1;
    return tz;
}

int64_t items_of(const vc_vdesc& d, int kind)
{
    switch (kind)
    {
    case VK_EZ: return cdiv((int64_t)d.nx * d.ny, kRowsZ);
    case VK_EY: return (int64_t)d.nx * cdiv(d.nz, d.tzy);
    case VK_EX: return (int64_t)d.ny * cdiv(d.nz, d.tzx);
    case VK_MS:
        return cdiv(d.nz, kMeasZ) * cdiv(d.ny, kMeasY) * cdiv(d.nx, kMeasXC);
    case VK_LI:
        return cdiv((int64_t)d.nx * d.ny * cdiv(d.nz, 512), kLiWaves);
    case VK_TT:
        return (int64_t)d.ntx * d.nty * d.nwz;
    case VK_LN:
        return cdiv((int64_t)d.nx * d.ny * d.nwz, kTpWords);
    default:
    {
        const int p = kind - VK_TP, sx = p & 1, sy = p >> 1;
        const int64_t nxh = (d.nx - sx + 1) / 2, nyh = (d.ny - sy + 1) / 2;
        return cdiv(nxh * nyh * d.nwz, kTpWords);
    }
    }
}

bool xcd_kind(int kind)
{
    return kind == VK_EY || kind == VK_EX || kind == VK_MS || kind == VK_TT;
}

void drop_graphs(vc_sched* s);

// captured launches hold the views by value: every change re-captures
int upload_desc(vc_ctx* ctx, vc_sched* s)
{
    // the staging buffer may still feed an earlier copy
    VC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    drop_graphs(s);
    std::memcpy(s->h_stage, s->h_desc.data(), sizeof(vc_vdesc) * s->n);
    VC_HIP(ctx, hipMemcpyAsync(s->d_desc, s->h_stage,
                               sizeof(vc_vdesc) * s->n,
                               hipMemcpyHostToDevice, ctx->stream));
    return VC_OK;
}

void drop_graphs(vc_sched* s)
{
    for (auto& g : s->graphs)
        (void)hipGraphExecDestroy(g.second);
    s->graphs.clear();
}

} // namespace

vc_schedule vc_schedule_default()
{
    vc_schedule sc;
    sc.edt_lds_kb = 16;
    sc.edt_grid_per_cu = 0;
    sc.thin_grid_per_cu = 0;
    sc.sweeps_per_sync = 0;
    sc.use_graph = 0;
    sc.tail_sweep = -1;
    sc.tile_sweep = 0;
    return sc;
}

int vc_vol_ensure(vc_ctx* ctx, vc_vol* v)
{
    if ((size_t)v->nx * 4 > 65536 || (size_t)v->ny * 2 > 65536)
        return vc_fail(ctx, VC_ERR_INVALID,
                       "axis too long for the LDS column tiles (nx <= 16384, "
                       "ny <= 32768)");
    const size_t N = (size_t)v->N;
    const size_t words = (size_t)v->nx * v->ny * ((v->nz + 63) / 64);
    int rc;
    if (!v->edt && (rc = vc_alloc(ctx, (void**)&v->edt, 4 * N)) != VC_OK)
        return rc;
    if (!v->ft && (rc = vc_alloc(ctx, (void**)&v->ft, 4 * N)) != VC_OK)
        return rc;
    if (!v->lam2 && (rc = vc_alloc(ctx, (void**)&v->lam2, 4 * N)) != VC_OK)
        return rc;
    if (!v->lvl && (rc = vc_alloc(ctx, (void**)&v->lvl, N)) != VC_OK)
        return rc;
    if (!v->label && (rc = vc_alloc(ctx, (void**)&v->label, N)) != VC_OK)
        return rc;
    if (!v->alive &&
        (rc = vc_alloc(ctx, (void**)&v->alive, 8 * words)) != VC_OK)
        return rc;
    if (!v->elig && (rc = vc_alloc(ctx, (void**)&v->elig, 8 * words)) != VC_OK)
        return rc;
    if (!v->dflag && (rc = vc_alloc(ctx, (void**)&v->dflag, 2 * words)) != VC_OK)
        return rc;
    if (!v->alive2 &&
        (rc = vc_alloc(ctx, (void**)&v->alive2, 8 * words)) != VC_OK)
        return rc;
    const size_t tiles = (size_t)cdiv(v->nx, kTileXY) * cdiv(v->ny, kTileXY) *
                         ((v->nz + 63) / 64);
    if (!v->tchg && (rc = vc_alloc(ctx, (void**)&v->tchg, 2 * 4 * tiles)) != VC_OK)
        return rc;
    if (!v->eflag)
    {
        const size_t nf = (size_t)v->nz * ((size_t)v->nx + v->ny);
        if ((rc = vc_alloc(ctx, (void**)&v->eflag, nf)) != VC_OK)
            return rc;
        VC_HIP(ctx, hipMemsetAsync(v->eflag, 0, nf, ctx->stream));
    }
    return VC_OK;
}

void vc_sched_free(vc_sched* s)
{
    if (!s)
        return;
    drop_graphs(s);
    vc_free(s->d_desc);
    for (int k = 0; k < VK_COUNT; ++k)
        vc_free(s->d_items[k]);
    vc_free(s->d_T);
    vc_free(s->d_L);
    vc_free(s->d_ctl);
    vc_free(s->d_gany);
    vc_free(s->d_vany);
    vc_free(s->d_ctr);
    if (s->h_stage)
        (void)hipHostFree(s->h_stage);
    if (s->h_lv)
        (void)hipHostFree(s->h_lv);
    if (s->h_ring)
        (void)hipHostFree(s->h_ring);
    if (s->h_ctr)
        (void)hipHostFree(s->h_ctr);
    delete s;
}

int vc_sched_apply(vc_ctx* ctx, vc_sched* s, const vc_schedule& sc)
{
    if (sc.edt_lds_kb != 8 && sc.edt_lds_kb != 16 && sc.edt_lds_kb != 32 &&
        sc.edt_lds_kb != 64)
        return vc_fail(ctx, VC_ERR_INVALID, "edt_lds_kb must be 8, 16, 32 or 64");
    if (sc.edt_grid_per_cu < 0 || sc.edt_grid_per_cu > 64 ||
        sc.thin_grid_per_cu < 0 || sc.thin_grid_per_cu > 64)
        return vc_fail(ctx, VC_ERR_INVALID, "grid_per_cu must be in [0, 64]");
    // the host reads a segment's flags from a ring of VC_RING entries whose
    // entries are cleared VC_RING / 2 sweeps ahead
    if (sc.sweeps_per_sync < -1 || sc.sweeps_per_sync > VC_RING / 2)
        return vc_fail(ctx, VC_ERR_INVALID,
                       "sweeps_per_sync must be in [-1, %d]", VC_RING / 2);
    if (sc.use_graph != 0 && sc.use_graph != 1)
        return vc_fail(ctx, VC_ERR_INVALID, "use_graph must be 0 or 1");
    if (sc.tail_sweep < -1 || sc.tail_sweep > 1000000)
        return vc_fail(ctx, VC_ERR_INVALID, "tail_sweep must be in [-1, 1e6]");
    if (sc.tile_sweep < -1 || sc.tile_sweep > 1000000)
        return vc_fail(ctx, VC_ERR_INVALID, "tile_sweep must be in [-1, 1e6]");
    VC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    drop_graphs(s);
    s->sc = sc;
    const size_t budget = (size_t)sc.edt_lds_kb * 1024;
    for (auto& d : s->h_desc)
    {
        // the y pass stages int16 but is sized like the x pass (4 bytes per
        // element): at 512^3 an 8-wide y tile (8 KB of a 16 KB budget, more
        // resident workgroups) runs 10% faster than a 16-wide one
        d.tzy = pick_tile(d.ny, 4, budget, d.nz);
        // the x pass gets twice the budget, at most 64 KB: 16-wide x tiles
        // (32 KB at 512^3, 64-byte row segments) measured 3.5 % faster for
        // the whole EDT than 8-wide ones
        d.tzx = pick_tile(d.nx, 4, std::min(2 * budget, (size_t)65536), d.nz);
    }
    std::memset(s->lds, 0, sizeof(s->lds));
    s->ez_stride = 0;
    for (const auto& d : s->h_desc)
    {
        s->ez_stride = std::max(s->ez_stride, d.nwz);
        s->lds[VK_EY] = std::max(s->lds[VK_EY],
                                 (size_t)vc_edt_ystride(d.tzy) * d.ny * 2);
        s->lds[VK_EX] = std::max(s->lds[VK_EX], (size_t)d.tzx * d.nx * 4);
    }
    s->lds[VK_EZ] = (size_t)kRowsZ * s->ez_stride * 8;
    for (int k = 0; k < VK_COUNT; ++k)
    {
        vc_free(s->d_items[k]);
        s->d_items[k] = nullptr;
        int64_t tot = 0;
        for (const auto& d : s->h_desc)
            tot += items_of(d, k);
        if (tot >= (int64_t(1) << 31) / 8)
            return vc_fail(ctx, VC_ERR_INVALID,
                           "batch too large: %lld work items of kind %d",
                           (long long)tot, k);
        s->total[k] = (uint32_t)tot;
        if (s->n == 1 || tot == 0)
            continue;
        std::vector<uint64_t> tab((size_t)tot);
        size_t o = 0;
        for (int v = 0; v < s->n; ++v)
        {
            const int64_t c = items_of(s->h_desc[v], k);
            for (int64_t j = 0; j < c; ++j)
                tab[o++] = ((uint64_t)v << 32) | (uint64_t)j;
        }
        int rc = vc_alloc(ctx, (void**)&s->d_items[k], 8 * tab.size());
        if (rc != VC_OK)
            return rc;
        VC_HIP(ctx, hipMemcpy(s->d_items[k], tab.data(), 8 * tab.size(),
                              hipMemcpyHostToDevice));
    }
    return upload_desc(ctx, s);
}

int vc_sched_build(vc_ctx* ctx, vc_vol* const* vols, int n,
                   const vc_schedule& sc, vc_sched** out)
{
    *out = nullptr;
    vc_sched* s = new (std::nothrow) vc_sched();
    if (!s)
        return VC_ERR_NOMEM;
    s->n = n;
    s->vols.assign(vols, vols + n);
    s->h_desc.resize(n);
    int rc = VC_OK;
    for (int i = 0; i < n && rc == VC_OK; ++i)
    {
        vc_vol* v = vols[i];
        rc = vc_vol_ensure(ctx, v);
        if (rc != VC_OK)
            break;
        vc_vdesc& d = s->h_desc[i];
        std::memset(&d, 0, sizeof(d));
        d.nx = v->nx;
        d.ny = v->ny;
        d.nz = v->nz;
        d.nwz = (v->nz + 63) / 64;
        d.pk = v->pk;
        d.mkind = v->measure_kind;
        d.occ = v->occ;
        d.edt = v->edt;
        d.ft = v->ft;
        d.lam2 = v->lam2;
        d.meas = v->meas_f;
        d.lvl = v->lvl;
        d.label = v->label;
        d.alive = v->alive;
        d.elig = v->elig;
        const size_t words = (size_t)v->nx * v->ny * d.nwz;
        d.d0 = v->dflag;
        d.d1 = v->dflag + words;
        d.alive2 = v->alive2;
        d.tchg = v->tchg;
        d.fy = v->eflag;
        d.fx = v->eflag + (size_t)v->nx * v->nz;
        d.ntx = (int)cdiv(v->nx, kTileXY);
        d.nty = (int)cdiv(v->ny, kTileXY);
        s->voxels += v->N;
        s->words += (int64_t)words;
    }
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&s->d_desc, sizeof(vc_vdesc) * n);
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&s->d_T, 8 * (size_t)n * VC_MAX_LEVELS);
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&s->d_L, 4 * (size_t)n * VC_MAX_LEVELS);
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&s->d_ctl, 4 * 64);
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&s->d_gany, 4 * VC_RING);
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&s->d_vany, 16 * (size_t)n);
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&s->d_ctr,
                      8 * (size_t)VC_CTR_SLOTS * VC_CTR_WORDS);
    if (rc == VC_OK &&
        (hipHostMalloc((void**)&s->h_stage, sizeof(vc_vdesc) * n,
                       hipHostMallocDefault) != hipSuccess ||
         hipHostMalloc(&s->h_lv, 12 * (size_t)n * VC_MAX_LEVELS,
                       hipHostMallocDefault) != hipSuccess ||
         hipHostMalloc((void**)&s->h_ring, 4 * VC_RING,
                       hipHostMallocDefault) != hipSuccess ||
         hipHostMalloc((void**)&s->h_ctr,
                       8 * (size_t)VC_CTR_SLOTS * VC_CTR_WORDS,
                       hipHostMallocDefault) != hipSuccess))
        rc = vc_fail(ctx, VC_ERR_NOMEM, "pinned host allocation failed");
    if (rc == VC_OK)
        rc = vc_sched_apply(ctx, s, sc);
    if (rc != VC_OK)
    {
        vc_sched_free(s);
        return rc;
    }
    *out = s;
    return VC_OK;
}

int vc_sched_sync_measure(vc_ctx* ctx, vc_sched* s)
{
    bool changed = false;
    for (int i = 0; i < s->n; ++i)
    {
        vc_vdesc& d = s->h_desc[i];
        const vc_vol* v = s->vols[i];
        if (d.mkind != v->measure_kind || d.meas != v->meas_f)
        {
            d.mkind = v->measure_kind;
            d.meas = v->meas_f;
            changed = true;
        }
    }
    return changed ? upload_desc(ctx, s) : VC_OK;
}

vc_work vc_sched_work(const vc_sched* s, int kind)
{
    vc_work w;
    w.D = s->d_desc;
    w.items = s->d_items[kind];
    w.total = s->total[kind];
    if (s->n == 1)
        w.one = s->h_desc[0];
    else
        std::memset(&w.one, 0, sizeof(w.one));
    return w;
}

unsigned vc_sched_grid(const vc_ctx* ctx, const vc_sched* s, int kind,
                       int per_cu)
{
    const int64_t tot = s->total[kind];
    if (tot == 0)
        return 0;
    // XCD-mapped kinds walk 8 * ceil(total / 8) slots
    const int64_t span = xcd_kind(kind) ? 8 * cdiv(tot, 8) : tot;
    if (per_cu <= 0)
        return (unsigned)span;
    const int64_t g = 8 * cdiv((int64_t)per_cu * std::max(ctx->cu_count, 1), 8);
    return (unsigned)std::min(span, g);
}

int vc_vol_sched(vc_ctx* ctx, vc_vol* v)
{
    if (v->self)
        return VC_OK;
    return vc_sched_build(ctx, &v, 1, vc_schedule_default(), &v->self);
}
