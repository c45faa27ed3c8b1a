// vc_thin.hip -- (26,6) simple-point table and measure-ordered subfield
// thinning, gfx950.
//
// Semantics (identical to the CPU oracle's dense schedule): for each level k
// (ascending), repeat sweeps until one removes nothing; a sweep visits the 8
// subfields s = (x&1)|(y&1)<<1|(z&1)<<2 in ascending order and removes, in
// parallel, every alive voxel of subfield s that is eligible at level k
// (lvl <= k) and (26,6)-simple in the current image.  Voxels of one subfield
// are never 26-adjacent, so the parallel removal equals any sequential order.
//
// Work-efficient execution: a voxel's test result can only change when a
// voxel of its 26-neighbourhood is removed, so instead of sweeping the volume
// every pass the kernels keep 16 work lists (8 subfields x sweep parity).  A
// level is seeded with the border voxels that became eligible at that level;
// every removal appends its alive, eligible, not-yet-queued neighbours to the
// list of the next pass of their subfield.  A level ends when every list is
// empty, which is exactly when a dense sweep would remove nothing.  The
// neighbour appends of a pass all target the same list per neighbour offset
// (the subfield of p + offset is s XOR parity(offset)), so they are
// aggregated per wavefront with one atomic per list.
#include <cstring>

#include "vc_internal.h"

namespace
{

constexpr int kBlock = 256;

// ------------------------------------------------------------ cube masks
struct CubeMasks
{
    uint32_t zlo, zhi, ylo, yhi, n6, n18, n26;
};

constexpr CubeMasks make_masks()
{
    CubeMasks m{0, 0, 0, 0, 0, 0, 0};
    for (int j = 0; j < 27; ++j)
    {
        int dz = j % 3 - 1, dy = (j / 3) % 3 - 1, dx = j / 9 - 1;
        int l1 = (dx != 0) + (dy != 0) + (dz != 0);
        if (dz == -1)
            m.zlo |= 1u << j;
        if (dz == 1)
            m.zhi |= 1u << j;
        if (dy == -1)
            m.ylo |= 1u << j;
        if (dy == 1)
            m.yhi |= 1u << j;
        if (l1 == 1)
            m.n6 |= 1u << j;
        if (l1 >= 1 && l1 <= 2)
            m.n18 |= 1u << j;
        if (l1 >= 1)
            m.n26 |= 1u << j;
    }
    return m;
}

constexpr CubeMasks kM = make_masks();
constexpr uint32_t kAll = 0x7FFFFFFu;

// 26-dilation inside the 3x3x3 cube (cell j = dz+1 + 3(dy+1) + 9(dx+1))
__device__ __forceinline__ uint32_t grow26(uint32_t m)
{
    uint32_t r = (m | ((m << 1) & ~kM.zlo) | ((m >> 1) & ~kM.zhi)) & kAll;
    r = (r | ((r << 3) & ~kM.ylo) | ((r >> 3) & ~kM.yhi)) & kAll;
    return (r | (r << 9) | (r >> 9)) & kAll;
}

__device__ __forceinline__ uint32_t grow6(uint32_t m)
{
    uint32_t r = m | ((m << 1) & ~kM.zlo) | ((m >> 1) & ~kM.zhi) |
                 ((m << 3) & ~kM.ylo) | ((m >> 3) & ~kM.yhi) | (m << 9) |
                 (m >> 9);
    return r & kAll;
}

// (26,6) simple point, Bertrand-Malandain: exactly one 26-component of the
// object in N26*(p) and exactly one 6-component of the background in N18*(p)
// 6-adjacent to p.  `code` = 26 neighbour bits, centre removed.
__device__ bool simple26(uint32_t code)
{
    const uint32_t X = ((code & 0x1FFFu) | (((code >> 13) & 0x1FFFu) << 14));
    if (!X)
        return false;
    uint32_t c = X & (0u - X), prev;
    do
    {
        prev = c;
        c = grow26(c) & X;
    } while (c != prev);
    if (c != X)
        return false;
    const uint32_t B = ~X & kM.n18;
    const uint32_t f6 = B & kM.n6;
    if (!f6)
        return false;
    c = f6 & (0u - f6);
    do
    {
        prev = c;
        c = grow6(c) & B;
    } while (c != prev);
    return (f6 & ~c) == 0;
}

__global__ void __launch_bounds__(kBlock) lut_build(uint32_t* lut)
{
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= (1u << 21))
        return;
    uint32_t word = 0;
    for (uint32_t b = 0; b < 32; ++b)
        word |= (uint32_t)simple26((w << 5) | b) << b;
    lut[w] = word;
}

// ------------------------------------------------------------ levels
struct LevelsU64
{
    uint64_t T[VC_MAX_LEVELS];
};
struct LevelsF32
{
    float L[VC_MAX_LEVELS];
};

__global__ void __launch_bounds__(kBlock)
    thin_init_exact(const uint8_t* __restrict__ occ,
                    const uint32_t* __restrict__ lam2, uint8_t* __restrict__ lvl,
                    uint8_t* __restrict__ label, int64_t N, LevelsU64 lv,
                    int nlev)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const bool in = occ[i] != 0;
    uint8_t l = 255;
    if (in)
    {
        const uint64_t m = lam2[i];
        for (int k = 0; k < nlev; ++k)
            if (m < lv.T[k])
            {
                l = (uint8_t)k;
                break;
            }
    }
    lvl[i] = l;
    label[i] = in ? 1 : 0;
}

__global__ void __launch_bounds__(kBlock)
    thin_init_f32(const uint8_t* __restrict__ occ,
                  const float* __restrict__ meas, uint8_t* __restrict__ lvl,
                  uint8_t* __restrict__ label, int64_t N, LevelsF32 lv,
                  int nlev)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const bool in = occ[i] != 0;
    uint8_t l = 255;
    if (in)
    {
        const float m = meas[i];
        for (int k = 0; k < nlev; ++k)
            if (m < lv.L[k])
            {
                l = (uint8_t)k;
                break;
            }
    }
    lvl[i] = l;
    label[i] = in ? 1 : 0;
}

__device__ __forceinline__ uint64_t lanes_below()
{
    const int lane = threadIdx.x & 63;
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// wavefront-aggregated append of `want` lanes' `val` to list `L`
__device__ __forceinline__ void wave_append(bool want, uint32_t val, int L,
                                            uint32_t* qcnt, uint32_t* qbuf,
                                            int64_t qcap)
{
    const uint64_t m = __ballot(want);
    if (!m)
        return;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if ((int)(threadIdx.x & 63) == leader)
        base = atomicAdd(&qcnt[L], (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (want)
        qbuf[(int64_t)L * qcap + base + (uint32_t)__popcll(m & lanes_below())] =
            val;
}

// seed level k: alive border voxels with lvl == k go to list (s, parity 0)
__global__ void __launch_bounds__(kBlock)
    thin_seed(const uint8_t* __restrict__ label, const uint8_t* __restrict__ lvl,
              uint32_t* __restrict__ qbits, uint32_t* __restrict__ qbuf,
              uint32_t* __restrict__ qcnt, int64_t qcap, int nx, int ny, int nz,
              vc_pack pk, int k)
{
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int x = blockIdx.z;
    bool want = false;
    const int64_t sy = nz, sx = (int64_t)ny * nz;
    const int64_t i = x * sx + y * sy + z;
    if (z < nz && label[i] == 1 && lvl[i] == k)
    {
        bool border = x == 0 || x == nx - 1 || y == 0 || y == ny - 1 ||
                      z == 0 || z == nz - 1;
        border = border || label[i - sx] != 1 || label[i + sx] != 1 ||
                 label[i - sy] != 1 || label[i + sy] != 1 ||
                 label[i - 1] != 1 || label[i + 1] != 1;
        if (border)
        {
            atomicOr(&qbits[i >> 5], 1u << (i & 31));
            want = true;
        }
    }
    const int s = (x & 1) | ((y & 1) << 1) | ((z & 1) << 2);
    const uint32_t val = vc_pack_xyz(pk, x, y, z);
    // the lanes of a wavefront share x and y; only the z parity differs
#pragma unroll
    for (int zb = 0; zb < 2; ++zb)
    {
        const int t = (x & 1) | ((y & 1) << 1) | (zb << 2);
        wave_append(want && s == t, val, t * 2, qcnt, qbuf, qcap);
    }
}

// one subfield pass: list (s, par) of level k
__global__ void __launch_bounds__(kBlock)
    thin_pass(uint8_t* __restrict__ label, const uint8_t* __restrict__ lvl,
              uint32_t* __restrict__ qbits, uint32_t* __restrict__ qbuf,
              uint32_t* __restrict__ qcnt, const uint32_t* __restrict__ lut,
              int64_t qcap, int nx, int ny, int nz, vc_pack pk, int s, int par,
              int k, int reset_list, unsigned long long* __restrict__ stats)
{
    const int L = s * 2 + par;
    const uint32_t n = qcnt[L];
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        if (reset_list >= 0)
            qcnt[reset_list] = 0;
        if (n)
            atomicAdd(&stats[2], 1ull);
    }
    const int64_t sy = nz, sx = (int64_t)ny * nz;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride)
    {
        const uint32_t idx = base + threadIdx.x;
        const bool valid = idx < n;
        int x = 0, y = 0, z = 0;
        int64_t i = 0;
        bool removed = false;
        if (valid)
        {
            vc_unpack_xyz(pk, qbuf[(int64_t)L * qcap + idx], x, y, z);
            i = x * sx + y * sy + z;
            atomicAnd(&qbits[i >> 5], ~(1u << (i & 31)));
            uint32_t code = 0;
            int b = 0;
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
                for (int dy = -1; dy <= 1; ++dy)
                {
                    const bool rin = (unsigned)(x + dx) < (unsigned)nx &&
                                     (unsigned)(y + dy) < (unsigned)ny;
                    const uint8_t* row = label + (i + dx * sx + dy * sy);
#pragma unroll
                    for (int dz = -1; dz <= 1; ++dz)
                    {
                        if (dx == 0 && dy == 0 && dz == 0)
                            continue;
                        const bool in = rin && (unsigned)(z + dz) < (unsigned)nz;
                        if (in && row[dz] == 1)
                            code |= 1u << b;
                        ++b;
                    }
                }
            removed = (lut[code >> 5] >> (code & 31)) & 1u;
            if (removed)
                label[i] = (uint8_t)(2 + k);
        }
        {
            const uint64_t mv = __ballot(valid);
            const uint64_t mr = __ballot(removed);
            if ((threadIdx.x & 63) == 0)
            {
                atomicAdd(&stats[0], (unsigned long long)__popcll(mv));
                if (mr)
                    atomicAdd(&stats[3 + k], (unsigned long long)__popcll(mr));
            }
        }
        // append the neighbours of removed voxels to their next pass
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                for (int dz = -1; dz <= 1; ++dz)
                {
                    if (dx == 0 && dy == 0 && dz == 0)
                        continue;
                    const int sq = s ^ ((dx & 1) | ((dy & 1) << 1) |
                                        ((dz & 1) << 2));
                    const int Lq = sq * 2 + (sq > s ? par : par ^ 1);
                    bool want = false;
                    uint32_t val = 0;
                    if (removed && (unsigned)(x + dx) < (unsigned)nx &&
                        (unsigned)(y + dy) < (unsigned)ny &&
                        (unsigned)(z + dz) < (unsigned)nz)
                    {
                        const int64_t j = i + dx * sx + dy * sy + dz;
                        if (label[j] == 1 && lvl[j] <= k)
                        {
                            const uint32_t bit = 1u << (j & 31);
                            const uint32_t old = atomicOr(&qbits[j >> 5], bit);
                            want = !(old & bit);
                            val = vc_pack_xyz(pk, x + dx, y + dy, z + dz);
                        }
                    }
                    wave_append(want, val, Lq, qcnt, qbuf, qcap);
                }
    }
}

} // namespace

int vc_lut_build(vc_ctx* ctx)
{
    int rc = vc_alloc(ctx, (void**)&ctx->d_lut, sizeof(uint32_t) << 21);
    if (rc != VC_OK)
        return rc;
    lut_build<<<(1u << 21) / kBlock, kBlock, 0, ctx->stream>>>(ctx->d_lut);
    VC_CHECK_LAUNCH(ctx);
    return VC_OK;
}

int vc_thin_run(vc_ctx* ctx, vc_vol* v, const float* levels, int n,
                vc_thin_stats* st)
{
    VC_HIP(ctx, hipSetDevice(ctx->device));
    const size_t N = (size_t)v->N;
    int rc;
    if (!v->label && (rc = vc_alloc(ctx, (void**)&v->label, N)) != VC_OK)
        return rc;
    if (!v->lvl && (rc = vc_alloc(ctx, (void**)&v->lvl, N)) != VC_OK)
        return rc;
    const size_t nwords = (N + 31) / 32;
    if (!v->qbits &&
        (rc = vc_alloc(ctx, (void**)&v->qbits, 4 * nwords)) != VC_OK)
        return rc;
    if (!v->qbuf)
    {
        v->qcap = (int64_t)((v->nx + 1) / 2) * ((v->ny + 1) / 2) *
                  ((v->nz + 1) / 2);
        if ((rc = vc_alloc(ctx, (void**)&v->qbuf,
                           sizeof(uint32_t) * 16 * (size_t)v->qcap)) != VC_OK)
            return rc;
    }
    if (!v->qcnt && (rc = vc_alloc(ctx, (void**)&v->qcnt, 64 * 4)) != VC_OK)
        return rc;
    if (!v->tstats &&
        (rc = vc_alloc(ctx, (void**)&v->tstats, 8 * (3 + VC_MAX_LEVELS))) !=
            VC_OK)
        return rc;
    uint32_t* hcnt = nullptr;
    VC_HIP(ctx, hipHostMalloc((void**)&hcnt, 64 * 4, hipHostMallocDefault));

    hipStream_t S = ctx->stream;
    hipError_t e = hipMemsetAsync(v->qbits, 0, 4 * nwords, S);
    if (e == hipSuccess)
        e = hipMemsetAsync(v->qcnt, 0, 64 * 4, S);
    if (e == hipSuccess)
        e = hipMemsetAsync(v->tstats, 0, 8 * (3 + VC_MAX_LEVELS), S);
    int64_t launches = 0;
    const unsigned nb = (unsigned)((N + kBlock - 1) / kBlock);
    if (e == hipSuccess)
    {
        if (v->measure_kind == 2)
        {
            LevelsF32 lv{};
            for (int k = 0; k < n; ++k)
                lv.L[k] = levels[k];
            thin_init_f32<<<nb, kBlock, 0, S>>>(v->occ, v->meas_f, v->lvl,
                                                 v->label, v->N, lv, n);
        }
        else
        {
            LevelsU64 lv{};
            for (int k = 0; k < n; ++k)
                lv.T[k] = (uint64_t)vc_level_threshold(levels[k]);
            thin_init_exact<<<nb, kBlock, 0, S>>>(v->occ, v->lam2, v->lvl,
                                                   v->label, v->N, lv, n);
        }
        launches++;
        e = hipGetLastError();
    }
    const dim3 gseed((unsigned)((v->nz + kBlock - 1) / kBlock), (unsigned)v->ny,
                     (unsigned)v->nx);
    const unsigned gpass = (unsigned)(ctx->cu_count > 0 ? ctx->cu_count * 4 : 1024);
    const int kBatch = 16; // two sweeps, keeps (s, parity) aligned per batch
    const int64_t kMaxPasses = int64_t(1) << 24;
    for (int k = 0; k < n && e == hipSuccess; ++k)
    {
        if (k > 0)
            e = hipMemsetAsync(v->qcnt, 0, 16 * 4, S);
        if (e != hipSuccess)
            break;
        thin_seed<<<gseed, kBlock, 0, S>>>(v->label, v->lvl, v->qbits, v->qbuf,
                                            v->qcnt, v->qcap, v->nx, v->ny,
                                            v->nz, v->pk, k);
        launches++;
        e = hipGetLastError();
        int64_t t = 0;
        int last = -1;
        while (e == hipSuccess)
        {
            for (int b = 0; b < kBatch && e == hipSuccess; ++b, ++t)
            {
                const int s = (int)(t & 7), par = (int)((t >> 3) & 1);
                thin_pass<<<gpass, kBlock, 0, S>>>(
                    v->label, v->lvl, v->qbits, v->qbuf, v->qcnt, ctx->d_lut,
                    v->qcap, v->nx, v->ny, v->nz, v->pk, s, par, k, last,
                    (unsigned long long*)v->tstats);
                launches++;
                e = hipGetLastError();
                last = s * 2 + par;
            }
            if (e != hipSuccess)
                break;
            e = hipMemcpyAsync(hcnt, v->qcnt, 16 * 4, hipMemcpyDeviceToHost, S);
            if (e == hipSuccess)
                e = hipStreamSynchronize(S);
            if (e != hipSuccess)
                break;
            bool done = true;
            for (int L = 0; L < 16; ++L)
                if (L != last && hcnt[L] != 0)
                    done = false;
            if (done)
                break;
            if (t > kMaxPasses)
            {
                (void)hipHostFree(hcnt);
                return vc_fail(ctx, VC_ERR_GUARD,
                               "thinning did not converge in %lld passes",
                               (long long)t);
            }
        }
    }
    if (e == hipSuccess && st)
    {
        uint64_t hs[3 + VC_MAX_LEVELS];
        e = hipMemcpyAsync(hs, v->tstats, sizeof(hs), hipMemcpyDeviceToHost, S);
        if (e == hipSuccess)
            e = hipStreamSynchronize(S);
        if (e == hipSuccess)
        {
            std::memset(st, 0, sizeof(*st));
            st->tests = (int64_t)hs[0];
            st->passes = (int64_t)hs[2];
            st->launches = launches;
            for (int k = 0; k < n; ++k)
                st->removed[k] = (int64_t)hs[3 + k];
        }
    }
    (void)hipHostFree(hcnt);
    if (e != hipSuccess)
        return vc_fail(ctx, VC_ERR_HIP, "thinning: %s", hipGetErrorString(e));
    v->thin_done = true;
    return VC_OK;
}
