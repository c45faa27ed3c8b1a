// vc_internal.h -- shared host/device declarations of libvoxelcore (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/vc_api.h"
#include "../../include/vc_batch.h"

// ---------------------------------------------------------------- context
struct vc_sched;

// scratch slots of a context (vc_ws_get); one per buffer of a call that
// uses them, no two live buffers share a slot
enum vc_ws_slot
{
    VC_WS_BND_TS,     // boundary: per-tile site counts
    VC_WS_BND_TF,     // boundary: per-tile face counts
    VC_WS_BND_CCNT,   // boundary: per-chunk counts
    VC_WS_BND_VMASK,  // boundary: per-voxel decision masks
    VC_WS_BND_CMAP,   // boundary: corner lattice -> site index
    VC_WS_BND_ACC,    // boundary: Euler accumulators and cube table
    VC_WS_BND_MARK,   // boundary: boundary marks
    VC_WS_BND_LAY8,   // boundary: thin-layer marks
    VC_WS_BND_LAB,    // boundary: thin-layer labels
    VC_WS_SCAN,       // chunk totals of the scans
    VC_WS_SLOTS
};

struct vc_ctx
{
    int device = 0;
    int cu_count = 0;
    hipStream_t stream = nullptr;
    // a second stream (default priority) for independent kernel chains within
    // one call, forked from and joined back to `stream` by the two events
    hipStream_t stream_aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    uint32_t* d_lut = nullptr; // simple-point table, 2^21 words
    std::string err;
    // autotuned schedules by batch shape (volumes, voxels, largest axis)
    std::map<std::vector<int64_t>, vc_schedule> tuned;
    // tile-sweep timing (vc_thin_timing): an event pair around each direct
    // thin_tile_sweep launch, or around each graph segment, read back at the
    // segment's host wait
    bool time_sweeps = false;
    int64_t sweep_launches = 0;
    double sweep_ms = 0.0;
    std::vector<hipEvent_t> ev_free;                         // reusable
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_open;  // not yet read
    std::vector<int> ev_open_n;                              // launches per pair
    // vc_skeleton_noalloc: this device's scratch array (device address) and
    // the host records of the call, built at context creation
    uint8_t* na_scratch = nullptr;
    vc_vol* na_vol = nullptr;
    vc_sched* na_sched = nullptr;
    // reusable device scratch by slot (vc_ws_get): grown on demand, kept until
    // the context is destroyed, so repeated calls allocate nothing
    void* ws[VC_WS_SLOTS] = {};
    size_t ws_bytes[VC_WS_SLOTS] = {};
};

// packed voxel coordinate: x << (by+bz) | y << bz | z, fits in 31 bits
struct vc_pack
{
    int bx, by, bz;
    uint32_t my, mz;
};

// LDS row (int16 elements) of the EDT y pass for TZ staged z columns: padded
// so that the 4 rows of a half-wavefront's 8 x 8 search block fall on
// disjoint LDS banks
inline __host__ __device__ int vc_edt_ystride(int tz)
{
    return tz + (tz >= 32 ? tz / 8 : 0);
}

// measure strip (z, y, x), one wavefront each: 62 computed z (lanes 1..62
// of a wavefront; lanes 0 and 63 hold the z halo), VC_MEAS_Y rows, 32 slices
#ifndef VC_MEAS_Y
#define VC_MEAS_Y 2
#endif
constexpr int kMeasZ = 62, kMeasY = VC_MEAS_Y, kMeasXC = 32;

struct vc_sched;

struct vc_vol
{
    vc_ctx* ctx = nullptr;
    int nx = 0, ny = 0, nz = 0;
    int64_t N = 0;
    vc_pack pk{};
    uint8_t* occ = nullptr;   // N, 0/1
    int32_t* edt = nullptr;   // N
    uint32_t* ft = nullptr;   // N, packed coordinate, 0xFFFFFFFF = none
    uint32_t* lam2 = nullptr; // N
    float* meas_f = nullptr;  // N, caller-supplied measure
    int measure_kind = 0;     // 0 none, 1 exact lam2, 2 float
    uint8_t* lvl = nullptr;   // N
    uint8_t* label = nullptr; // N
    // thinning state: bit planes packed along z (64 voxels per word)
    uint64_t* alive = nullptr; // alive voxels
    uint64_t* elig = nullptr;  // eligible at the current level (lvl <= k)
    uint8_t* dflag = nullptr;  // 2 x words: revisit flags per z parity
    uint64_t* alive2 = nullptr; // tile sweeps: the other alive plane
    uint32_t* tchg = nullptr;   // tile sweeps: 2 x tiles neighbour masks
    uint8_t* eflag = nullptr;   // EDT thick-tile flags: nx*nz (y pass), ny*nz (x pass)
    // boundary stage results (computed once per volume)
    bool bnd_done = false;
    float* bnd_sites = nullptr;   // 3 floats per site
    int32_t* bnd_quads = nullptr; // 4 site indices per quad
    int64_t bnd_nsites = 0, bnd_nquads = 0;
    vc_euler bnd_euler{};
    bool edt_done = false;
    bool thin_done = false;
    vc_sched* self = nullptr; // schedule of this volume alone
};

static constexpr uint32_t VC_NONE = 0xFFFFFFFFu;

// ---------------------------------------------------------------- errors
int vc_fail(vc_ctx* ctx, int code, const char* fmt, ...);

#define VC_HIP(ctx, call)                                                       \
    do                                                                          \
    {                                                                           \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess)                                                   \
            return vc_fail((ctx), VC_ERR_HIP, "%s: %s (%s:%d)", #call,          \
                           hipGetErrorString(e_), __FILE__, __LINE__);          \
    } while (0)

#define VC_CHECK_LAUNCH(ctx) VC_HIP(ctx, hipGetLastError())

int vc_alloc(vc_ctx* ctx, void** p, size_t bytes);
void vc_free(void* p);
// device scratch of slot `slot` with at least `bytes` bytes, kept by the
// context: contents are not preserved when it grows (the old buffer is freed
// after the device finishes with it); calls on one context run one at a time
int vc_ws_get(vc_ctx* ctx, int slot, size_t bytes, void** p);

// ---------------------------------------------------------------- schedules
// Every volume kernel walks a work table: item b of a launch is one tile of
// one volume.  A batch of volumes is one launch per pass over all of its
// tiles; a single volume is a batch of one (no table, item = local tile).

// per-volume view read by the kernels (device copy in the schedule)
struct vc_vdesc
{
    int nx, ny, nz, nwz;
    vc_pack pk;
    int tzy, tzx; // EDT column tiles: z columns per workgroup (passes y, x)
    int mkind;    // 1 exact lam2, 2 caller float measure
    int pad_;
    const uint8_t* occ;
    int32_t* edt;
    uint32_t* ft;
    uint32_t* lam2;
    const float* meas;
    uint8_t* lvl;
    uint8_t* label;
    uint64_t* alive;
    uint64_t* elig;
    uint8_t* d0; // revisit flags, z parity 0
    uint8_t* d1; // revisit flags, z parity 1
    uint64_t* alive2; // second alive plane of the tile sweeps
    uint32_t* tchg;   // 2 x tiles (by sweep parity): bit (dx+1)*9+(dy+1)*3+
                      // (dz+1) = the tile removed a voxel in the halo of its
                      // neighbour at (dx, dy, dz) (bit 13: anywhere)
    uint8_t* fy;      // EDT y-pass tiles to compute by divide and conquer
    uint8_t* fx;      // (set by the pass before, cleared by the tile)
    int ntx, nty;     // tiles along x and y (32 rows each)
};

// kernels with a work table
enum vc_kind
{
    VK_EZ = 0, // EDT pass z: 4 rows per item
    VK_EY,     // EDT pass y: one (x, z-tile) column tile per item
    VK_EX,     // EDT pass x: one (y, z-tile) column tile per item
    VK_MS,     // measure: one kMeasZ x kMeasY x kMeasXC wavefront strip per item
    VK_LI,     // level set-up: 4 wavefronts (512 z each) per item
    VK_TP,     // subfield pass, (x, y) parity 0..3: 256 words per item
    VK_TT = VK_TP + 4, // tile sweep: one 32 x 32 x 64 (x, y, z) tile per item
    VK_LN,             // level set-up for k > 0: 256 plane words per item
    VK_COUNT
};

// work table of one kind as the kernels see it
struct vc_work
{
    const vc_vdesc* D;
    const uint64_t* items; // volume << 32 | local item; null for one volume
    uint32_t total;
    vc_vdesc one; // the volume itself when items is null (kernel argument:
                  // no dependent load before the first data access)
};

// sweeps per segment when vc_schedule.sweeps_per_sync is 0 (adaptive)
constexpr int VC_SEG_MIN = 4, VC_SEG_MAX = 16;
// ring of per-sweep "removed anything" flags
constexpr int VC_RING = 64;
// stats counters: kSlots cache lines of [0] tests, [8 + k] removed at level k
constexpr int VC_CTR_SLOTS = 64;
constexpr int VC_CTR_WORDS = 8 + VC_MAX_LEVELS;

struct vc_sched
{
    int n = 0;
    std::vector<vc_vol*> vols;
    std::vector<vc_vdesc> h_desc;
    vc_vdesc* d_desc = nullptr;
    vc_vdesc* h_stage = nullptr; // pinned staging of h_desc
    uint64_t* d_items[VK_COUNT] = {};
    uint32_t total[VK_COUNT] = {};
    size_t lds[VK_COUNT] = {};
    int ez_stride = 0; // LDS words per wavefront in pass z
    int64_t voxels = 0, words = 0;
    vc_schedule sc{};
    // per-volume level thresholds, n x VC_MAX_LEVELS each
    uint64_t* d_T = nullptr;
    float* d_L = nullptr;
    void* h_lv = nullptr; // pinned staging of d_T then d_L
    std::vector<float> lv_cache;
    int lv_n = 0;
    // thinning control: [0] first sweep of the running segment, [1] level
    uint32_t* d_ctl = nullptr;
    uint32_t* d_gany = nullptr;  // VC_RING: sweep removed anything (batch)
    uint32_t* d_vany = nullptr;  // 4 per volume: ring of the same per volume
    unsigned long long* d_ctr = nullptr;
    uint32_t* h_ring = nullptr;  // pinned copy of d_gany
    unsigned long long* h_ctr = nullptr;
    std::map<int, hipGraphExec_t> graphs; // captured segments by length
};

int vc_sched_build(vc_ctx* ctx, vc_vol* const* vols, int n,
                   const vc_schedule& sc, vc_sched** out);
void vc_sched_free(vc_sched* s);
// rebuild the tables that depend on the schedule's launch parameters
int vc_sched_apply(vc_ctx* ctx, vc_sched* s, const vc_schedule& sc);
// refresh the measure kind / pointer of each volume in the device copy
int vc_sched_sync_measure(vc_ctx* ctx, vc_sched* s);
vc_work vc_sched_work(const vc_sched* s, int kind);
unsigned vc_sched_grid(const vc_ctx* ctx, const vc_sched* s, int kind,
                       int per_cu);
vc_schedule vc_schedule_default();
int vc_vol_ensure(vc_ctx* ctx, vc_vol* v); // allocate the path's buffers
int vc_vol_sched(vc_ctx* ctx, vc_vol* v);  // build v->self if missing

// ---------------------------------------------------------------- stages
int vc_lut_build(vc_ctx* ctx);
// vc_skeleton_noalloc set-up at context creation / release at destruction
int vc_noalloc_setup(vc_ctx* ctx);
void vc_noalloc_teardown(vc_ctx* ctx);
int vc_edt_sched(vc_ctx* ctx, vc_sched* s);
// n bytes of device memory set to v by a kernel on the context stream (no
// memset node: callers may be under graph capture)
int vc_fill_u8(vc_ctx* ctx, uint8_t* p, size_t n, uint8_t v);
int vc_measure_sched(vc_ctx* ctx, vc_sched* s);
// levels: n x n_levels (row i = ascending levels of volume i)
int vc_thin_sched(vc_ctx* ctx, vc_sched* s, const float* levels,
                  int n_levels, vc_thin_stats* st);
int vc_boundary_run(vc_ctx* ctx, vc_vol* vol, float* sites, int64_t site_cap,
                    int64_t* n_sites, int32_t* quads, int64_t quad_cap,
                    int64_t* n_quads, vc_euler* eu);
int vc_ft_to_linear(vc_ctx* ctx, vc_vol* vol, int32_t* d_dst);
// exclusive scan of n uint32 in device memory, in place, on the context
// stream (waits)
int vc_scan_u32(vc_ctx* ctx, uint32_t* a, int64_t n);
// cellcomplex::finalize: host-side validation of array inputs, and the
// device computation on device arrays (vc_voro.hip)
int vc_cc_check_host(vc_ctx* ctx, int32_t nv, int32_t ne_in,
                     const int32_t* edges_in, int32_t nf,
                     const int64_t* face_off, const int32_t* face_vts);
int vc_cc_finalize_dev(vc_ctx* ctx, int32_t ne_in, const int32_t* dE, int32_t nf,
                       const int64_t* dfo, const int32_t* dfv, int64_t nfv,
                       int32_t* deo, int32_t* dfe, int32_t* ne_out);

// host helpers
int64_t vc_level_threshold(float level);
// n in [1, VC_MAX_LEVELS], no NaN, ascending
int vc_check_levels(vc_ctx* ctx, const float* levels, int n);

__host__ __device__ inline uint32_t vc_pack_xyz(const vc_pack& p, int x, int y,
                                                int z)
{
    return ((uint32_t)x << (p.by + p.bz)) | ((uint32_t)y << p.bz) | (uint32_t)z;
}
__host__ __device__ inline void vc_unpack_xyz(const vc_pack& p, uint32_t v,
                                              int& x, int& y, int& z)
{
    z = (int)(v & p.mz);
    y = (int)((v >> p.bz) & p.my);
    x = (int)(v >> (p.by + p.bz));
}

// ---------------------------------------------------------------- work items
// Kernels are instantiated twice: kB = false for one volume (the view is the
// kernel-argument copy W.one, no table) and kB = true for a batch (table and
// views read through the constant address space, so that the uniform loads
// are scalar and the fields land in scalar registers).
typedef const __attribute__((address_space(4))) uint32_t* vc_cu32;
typedef const __attribute__((address_space(4))) uint64_t* vc_citems;
static_assert(sizeof(vc_vdesc) % 4 == 0, "vc_vdesc is copied as words");

template <bool kB>
__device__ __forceinline__ vc_vdesc vc_desc(const vc_work& W, int v)
{
    if (kB)
    {
        vc_vdesc out;
        const vc_cu32 src = (vc_cu32)(const void*)(W.D + v);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
        for (int i = 0; i < (int)(sizeof(vc_vdesc) / 4); ++i)
            dst[i] = src[i];
        return out;
    }
    return W.one;
}

// item b of a work table -> (volume, local tile)
template <bool kB>
__device__ __forceinline__ void vc_item(const vc_work& W, uint32_t b, int& v,
                                        uint32_t& lt)
{
    if (kB)
    {
        const uint64_t e = ((vc_citems)W.items)[b];
        v = (int)(e >> 32);
        lt = (uint32_t)e;
    }
    else
    {
        v = 0;
        lt = b;
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  XCD-mapped launches give consecutive items (which share cache lines
// and halos) to workgroups of one XCD: slot L -> item (L % 8) * per + L / 8,
// per = ceil(total / 8); with a grid that is a multiple of 8, every slot a
// workgroup visits stays on its XCD.
__device__ __forceinline__ uint32_t vc_xcd_item(uint32_t L, uint32_t per)
{
    return (L & 7u) * per + (L >> 3);
}

// ---- the Voronoi-side steps on device arrays (vc_voro.hip,
// vc_voro_merge.hip) and resident complexes built from device arrays
// (vc_ccthin_dev.hip): the pieces of the resident voxel-core chain

// the discrete Voronoi diagram of a feature transform in device arrays
// (vc_alloc'ed; freed by vc_voro_dev_free)
struct vc_voro_dev
{
    int64_t nv = 0, nf = 0, ns = 0;
    float* vts = nullptr;     // 3 nv
    int32_t* faces = nullptr; // 4 nf (quads)
    int32_t* fs = nullptr;    // 2 nf
    float* sites = nullptr;   // 3 ns
};
void vc_voro_dev_free(vc_voro_dev* d);
// vc_voro_from_ft into device arrays; waits
int vc_voro_from_ft_dev(vc_ctx* ctx, vc_vol* vol, vc_voro_dev* out);
// vc_voro_tag_points / vc_voro_radii / vc_voro_measures on device arrays,
// queued on the context stream (no wait); radii needs an nv-int scratch,
// measures take the adjacency of a resident complex (uint32 offsets)
int vc_voro_tag_dev(vc_ctx* ctx, vc_vol* vol, const float* pts, int64_t n,
                    uint8_t* valid);
int vc_voro_radii_dev(vc_ctx* ctx, int32_t nv, int32_t nf, const float* vts,
                      const int64_t* face_off, const int32_t* face_vts,
                      const float* sites, const int32_t* face_sites, int32_t* last,
                      float* r);
int vc_voro_measures_dev(vc_ctx* ctx, int32_t nv, int32_t ne, int32_t nf,
                         const float* sites, const int32_t* face_sites,
                         const uint8_t* face_valid, const uint32_t* ve_off,
                         const int32_t* ve, const uint32_t* ef_off, const int32_t* ef,
                         float* lam_v, float* lam_e, float* lam_f);
// vc_voro_extract_inside on device arrays: outputs vc_alloc'ed (freed by
// vc_voro_xi_free): vi (nv), eo (2 ne), em (ne), tr (3 nt), ff, fm (nt); waits
struct vc_voro_xi
{
    int64_t nv = 0, ne = 0, nt = 0;
    int32_t* vi = nullptr;
    int32_t* eo = nullptr;
    float* em = nullptr;
    int32_t* tr = nullptr;
    int32_t* ff = nullptr;
    float* fm = nullptr;
};
void vc_voro_xi_free(vc_voro_xi* x);
int vc_voro_extract_dev(vc_ctx* ctx, int32_t nv, int32_t ne, int32_t nf,
                        const int32_t* dE, const int64_t* dfo, const int32_t* dfv,
                        const uint8_t* dval, const float* dle, const float* dlf,
                        vc_voro_xi* out);
// the device path of vc_voro_merge_close_device on device arrays: *merges = 1
// (nothing written) when some edge joins two vertices within eps; otherwise
// old_to_new (o2n), the seeds (nvts, n_new), the kept faces' new polygons
// (nfo / nfv, nf_kept faces, nfv_new ids), remove / keep flags, each face's
// exclusive kept-face position (fpos, nf + 1) and the sorted new edges
// (edges, ne_new); vc_alloc'ed, freed by vc_voro_mg_free; waits
struct vc_voro_mg
{
    int32_t n_new = 0, nf_kept = 0;
    int64_t nfv_new = 0, ne_new = 0;
    int32_t* o2n = nullptr;
    int32_t* nvts = nullptr;
    int64_t* nfo = nullptr;
    int32_t* nfv = nullptr;
    uint8_t* keep = nullptr;
    uint8_t* rem = nullptr;
    uint32_t* fpos = nullptr;
    int32_t* edges = nullptr;
};
void vc_voro_mg_free(vc_voro_mg* m);
int vc_voro_merge_dev(vc_ctx* ctx, int32_t nv, int32_t ne, int32_t nf, const float* dv,
                      const int32_t* dE, const int64_t* dfo, const int32_t* dfv,
                      int64_t nfv, const uint8_t* dval, float eps, int only_inside,
                      vc_voro_mg* out, int* merges);

struct vc_cc;
// vc_cc_create from device arrays (copied; the inputs may be freed after)
int vc_cc_create_dev(vc_ctx* ctx, int32_t nv, int32_t ne_in, const int32_t* dE,
                     int32_t nf, const int64_t* dfo, const int32_t* dfv, int64_t nfv,
                     struct vc_cc** out);
// the device arrays of a resident complex
struct vc_cc_dview
{
    int32_t nv, ne, nf;
    int64_t nfe;
    const int32_t* E;
    const int64_t* fo;
    const int32_t* fv;
    const int32_t* fe;
    const uint32_t* ve_off;
    const int32_t* ve;
    const uint32_t* ef_off;
    const int32_t* ef;
};
void vc_cc_dview_get(const struct vc_cc* c, vc_cc_dview* v);
