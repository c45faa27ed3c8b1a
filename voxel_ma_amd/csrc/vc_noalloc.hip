// vc_noalloc.hip -- vc_skeleton_noalloc: EDT + measure + thinning of one
// volume held in caller device memory, with no device allocation.
//
// Every intermediate of the path (squared EDT, feature transform, lambda^2,
// level bytes, the thinning bit planes, revisit flags, tile neighbour masks,
// the per-volume view and the thinning control words) lives in one
// file-scope device array, `scratch`, of VC_NOALLOC_SCRATCH_BYTES.  Each GPU
// the library runs on holds its own copy at its own address until the
// process ends; the address is taken once per device with
// hipGetSymbolAddress at context creation and passed to the kernels as an
// ordinary pointer (the array's name is only a host-side stand-in).
//
// Every call on a device shares that device's array and the array keeps the
// last call's contents: the kernels initialise every word they read before
// reading it (the level set-up clears the control words and rings and
// writes the planes, the tile-sweep entry writes the neighbour masks, the
// statistics counters are cleared at the start of the thinning), and the
// level thresholds and the view are uploaded on every call.  Calls that
// overlap on one device (two host threads, two contexts of one device)
// race on the array.
//
// The host side of the call reuses a volume record and a one-volume schedule
// built, with their pinned staging buffers, by vc_ctx_create; the call
// itself makes no hipMalloc, hipHostMalloc or hipFree.
#include <cstring>
#include <mutex>
#include <new>

#include "vc_internal.h"

static __device__ __align__(256) uint8_t scratch[VC_NOALLOC_SCRATCH_BYTES];

namespace
{

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

int bits_for(int n)
{
    int b = 0;
    while ((1 << b) < n)
        b++;
    return b;
}

// byte offsets of the buffers in the scratch array for an nx x ny x nz
// volume; returns the bytes needed
struct Layout
{
    size_t edt, ft, lam2, lvl, alive, elig, dflag, alive2, tchg, eflag;
    size_t desc, T, L, ctl, gany, vany, ctr, end;
};

Layout layout(int nx, int ny, int nz)
{
    const size_t N = (size_t)nx * ny * nz;
    const size_t words = (size_t)nx * ny * ((nz + 63) / 64);
    const size_t tiles = (size_t)((nx + 31) / 32) * ((ny + 31) / 32) *
                         ((nz + 63) / 64);
    Layout l{};
    size_t o = 0;
    auto take = [&](size_t& at, size_t bytes) {
        at = o;
        o += up256(bytes);
    };
    take(l.edt, 4 * N);
    take(l.ft, 4 * N);
    take(l.lam2, 4 * N);
    take(l.lvl, N);
    take(l.alive, 8 * words);
    take(l.elig, 8 * words);
    take(l.dflag, 2 * words);
    take(l.alive2, 8 * words);
    take(l.tchg, 2 * 4 * tiles);
    take(l.eflag, (size_t)nz * ((size_t)nx + ny));
    take(l.desc, sizeof(vc_vdesc));
    take(l.T, 8 * (size_t)VC_MAX_LEVELS);
    take(l.L, 4 * (size_t)VC_MAX_LEVELS);
    take(l.ctl, 4 * 64);
    take(l.gany, 4 * (size_t)VC_RING);
    take(l.vany, 16);
    take(l.ctr, 8 * (size_t)VC_CTR_SLOTS * VC_CTR_WORDS);
    l.end = o;
    return l;
}

} // namespace

// At context creation, with the context's device current: the device
// address of this device's scratch array (taken once per device), and the
// host records of the call with their pinned staging buffers.
int vc_noalloc_setup(vc_ctx* ctx)
{
    if (ctx->device < 0 || ctx->device >= kMaxDevices)
        return vc_fail(ctx, VC_ERR_INVALID, "device %d beyond %d", ctx->device,
                       kMaxDevices);
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (!g_ptr[ctx->device])
        {
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
            {
                (void)hipGetLastError();
                return vc_fail(ctx, VC_ERR_HIP,
                               "hipGetSymbolAddress(scratch) on device %d: %s",
                               ctx->device, hipGetErrorString(e));
            }
            g_ptr[ctx->device] = p;
        }
        ctx->na_scratch = static_cast<uint8_t*>(g_ptr[ctx->device]);
    }
    vc_vol* v = new (std::nothrow) vc_vol();
    vc_sched* s = new (std::nothrow) vc_sched();
    if (!v || !s)
    {
        delete v;
        delete s;
        return VC_ERR_NOMEM;
    }
    v->ctx = ctx;
    s->n = 1;
    s->vols.assign(1, v);
    s->h_desc.resize(1);
    s->lv_cache.reserve(VC_MAX_LEVELS);
    ctx->na_vol = v;
    ctx->na_sched = s;
    if (hipHostMalloc((void**)&s->h_stage, sizeof(vc_vdesc),
                      hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&s->h_lv, 12 * (size_t)VC_MAX_LEVELS,
                      hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&s->h_ring, 4 * VC_RING, hipHostMallocDefault) !=
            hipSuccess ||
        hipHostMalloc((void**)&s->h_ctr, 8 * (size_t)VC_CTR_SLOTS * VC_CTR_WORDS,
                      hipHostMallocDefault) != hipSuccess)
    {
        (void)hipGetLastError();
        return vc_fail(ctx, VC_ERR_NOMEM, "pinned host allocation failed");
    }
    return VC_OK;
}

// releases the host records; the scratch array stays with the device
void vc_noalloc_teardown(vc_ctx* ctx)
{
    vc_sched* s = ctx->na_sched;
    if (s)
    {
        if (s->h_stage)
            (void)hipHostFree(s->h_stage);
        if (s->h_lv)
            (void)hipHostFree(s->h_lv);
        if (s->h_ring)
            (void)hipHostFree(s->h_ring);
        if (s->h_ctr)
            (void)hipHostFree(s->h_ctr);
        delete s; // its device pointers are scratch: never freed
    }
    delete ctx->na_vol;
    ctx->na_sched = nullptr;
    ctx->na_vol = nullptr;
}

extern "C" int vc_skeleton_noalloc(vc_ctx* ctx, const uint8_t* d_occ, int nx,
                                   int ny, int nz, const float* levels,
                                   int n_levels, uint8_t* d_label,
                                   vc_thin_stats* stats)
{
    if (!ctx || !d_occ || !d_label || !levels)
        return VC_ERR_INVALID;
    if (!ctx->na_scratch || !ctx->na_sched || !ctx->na_vol)
        return vc_fail(ctx, VC_ERR_STATE, "vc_skeleton_noalloc: not set up");
    int rc = vc_check_levels(ctx, levels, n_levels);
    if (rc != VC_OK)
        return rc;
    if (nx <= 0 || ny <= 0 || nz <= 0 || nx > 16384 || ny > 32768 ||
        nz > VC_MAX_AXIS)
        return vc_fail(ctx, VC_ERR_INVALID,
                       "vc_skeleton_noalloc: dims %d x %d x %d out of range "
                       "(nx <= 16384, ny <= 32768, nz <= %d)",
                       nx, ny, nz, VC_MAX_AXIS);
    const int64_t N = (int64_t)nx * ny * nz;
    const int bx = bits_for(nx), by = bits_for(ny), bz = bits_for(nz);
    const double maxd2 = (double)(nx - 1) * (nx - 1) +
                         (double)(ny - 1) * (ny - 1) + (double)(nz - 1) * (nz - 1);
    if (N > VC_NOALLOC_MAX_VOXELS || bx + by + bz > 31 || maxd2 > 2147483647.0)
        return vc_fail(ctx, VC_ERR_INVALID,
                       "vc_skeleton_noalloc: %d x %d x %d = %lld voxels exceeds "
                       "the limit of %lld voxels",
                       nx, ny, nz, (long long)N,
                       (long long)VC_NOALLOC_MAX_VOXELS);
    const Layout L = layout(nx, ny, nz);
    if (L.end > (size_t)VC_NOALLOC_SCRATCH_BYTES)
        return vc_fail(ctx, VC_ERR_INVALID,
                       "vc_skeleton_noalloc: %d x %d x %d needs %zu bytes of "
                       "scratch, more than the %lld available",
                       nx, ny, nz, L.end, (long long)VC_NOALLOC_SCRATCH_BYTES);
    VC_HIP(ctx, hipSetDevice(ctx->device));
    uint8_t* const base = ctx->na_scratch;
    vc_vol* v = ctx->na_vol;
    v->nx = nx;
    v->ny = ny;
    v->nz = nz;
    v->N = N;
    v->pk.bx = bx;
    v->pk.by = by;
    v->pk.bz = bz;
    v->pk.my = (1u << by) - 1u;
    v->pk.mz = (1u << bz) - 1u;
    v->occ = const_cast<uint8_t*>(d_occ); // read only
    v->edt = reinterpret_cast<int32_t*>(base + L.edt);
    v->ft = reinterpret_cast<uint32_t*>(base + L.ft);
    v->lam2 = reinterpret_cast<uint32_t*>(base + L.lam2);
    v->meas_f = nullptr;
    v->measure_kind = 0;
    v->lvl = base + L.lvl;
    v->label = d_label;
    v->alive = reinterpret_cast<uint64_t*>(base + L.alive);
    v->elig = reinterpret_cast<uint64_t*>(base + L.elig);
    v->dflag = base + L.dflag;
    v->alive2 = reinterpret_cast<uint64_t*>(base + L.alive2);
    v->tchg = reinterpret_cast<uint32_t*>(base + L.tchg);
    v->eflag = base + L.eflag;
    v->edt_done = v->thin_done = false;

    vc_sched* s = ctx->na_sched;
    vc_vdesc& d = s->h_desc[0];
    std::memset(&d, 0, sizeof(d));
    d.nx = nx;
    d.ny = ny;
    d.nz = nz;
    d.nwz = (nz + 63) / 64;
    d.pk = v->pk;
    d.mkind = 0;
    d.occ = v->occ;
    d.edt = v->edt;
    d.ft = v->ft;
    d.lam2 = v->lam2;
    d.meas = nullptr;
    d.lvl = v->lvl;
    d.label = v->label;
    d.alive = v->alive;
    d.elig = v->elig;
    const size_t words = (size_t)nx * ny * d.nwz;
    d.d0 = v->dflag;
    d.d1 = v->dflag + words;
    d.alive2 = v->alive2;
    d.tchg = v->tchg;
    d.fy = v->eflag;
    d.fx = v->eflag + (size_t)nx * nz;
    d.ntx = (nx + 31) / 32;
    d.nty = (ny + 31) / 32;
    s->voxels = N;
    s->words = (int64_t)words;
    s->d_desc = reinterpret_cast<vc_vdesc*>(base + L.desc);
    s->d_T = reinterpret_cast<uint64_t*>(base + L.T);
    s->d_L = reinterpret_cast<float*>(base + L.L);
    s->d_ctl = reinterpret_cast<uint32_t*>(base + L.ctl);
    s->d_gany = reinterpret_cast<uint32_t*>(base + L.gany);
    s->d_vany = reinterpret_cast<uint32_t*>(base + L.vany);
    s->d_ctr = reinterpret_cast<unsigned long long*>(base + L.ctr);
    // thresholds and view are uploaded on every call: another context of
    // this device may have used the array since
    s->lv_cache.clear();
    s->lv_n = 0;
    // one volume: no work tables, so this allocates nothing; it uploads the
    // view into the scratch array
    rc = vc_sched_apply(ctx, s, vc_schedule_default());
    // the EDT's thick-tile flags start cleared (the array may hold another
    // call's data there; a stale flag would cost time only)
    if (rc == VC_OK)
        rc = vc_fill_u8(ctx, v->eflag, (size_t)nz * ((size_t)nx + ny), 0);
    if (rc == VC_OK)
        rc = vc_edt_sched(ctx, s);
    if (rc == VC_OK)
        rc = vc_measure_sched(ctx, s);
    if (rc == VC_OK)
        rc = vc_thin_sched(ctx, s, levels, n_levels, stats);
    return rc;
}
