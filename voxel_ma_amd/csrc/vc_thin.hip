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
// Execution: the alive and the eligible (lvl <= k) voxels are bit planes
// packed along z, 64 voxels per word.  A pass is one launch over the words of
// the rows of its (x, y) parity, one thread per word: candidates of the pass's
// z parity are filtered with word operations (alive, eligible, at least one
// background face neighbour), and only those go through the simple-point
// table.  The thread owns its word for the pass, so removals need no atomics.
// Per-word revisit flags (one byte per word and z parity, plain stores) skip
// words whose neighbourhood did not change since their last visit.  A level
// ends after a sweep that removed nothing (later passes exit at once); the
// host checks the per-sweep "removed anything" flags once per batch of sweeps.
#include <cstring>
#include <vector>

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

// ------------------------------------------------------------ bit planes
// Alive and eligible voxels are bit planes packed along z: word (x, y, w)
// holds z in [64w, 64w+64), bit z & 63; index (x*ny + y)*nwz + w.
struct Plane
{
    int nx, ny, nz, nwz;
};

__device__ __forceinline__ int64_t widx(const Plane& P, int x, int y, int w)
{
    return ((int64_t)x * P.ny + y) * P.nwz + w;
}

// Level set-up.  A wavefront covers 512 consecutive z of one row (8 words);
// lane l handles z = z0 + 8l .. z0 + 8l + 7 and lanes 8j .. 8j+7 assemble
// word j.  Words holding an alive voxel that becomes eligible at level k are
// flagged for both z parities (every other eligible voxel was found
// non-simple when the previous level converged and its neighbourhood has not
// changed since).
template <int kMode> // 0: exact measure, 1: float measure, 2: later level
__global__ void __launch_bounds__(kBlock)
    thin_level_init(const uint8_t* __restrict__ occ,
                    const uint32_t* __restrict__ lam2,
                    const float* __restrict__ meas, uint8_t* __restrict__ lvl,
                    uint8_t* __restrict__ label, uint64_t* __restrict__ A,
                    uint64_t* __restrict__ E, uint8_t* __restrict__ D0,
                    uint8_t* __restrict__ D1, Plane P, int nchunk,
                    LevelsU64 lt, LevelsF32 lf, int nlev, int k)
{
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t rowi = wave / nchunk;
    if (rowi >= (int64_t)P.nx * P.ny)
        return;
    const int z0 = (int)(wave % nchunk) * 512 + 8 * lane;
    const int64_t base = rowi * P.nz;
    uint32_t am = 0, em = 0, nm = 0;
    uint8_t lv8[8], lb8[8];
#pragma unroll
    for (int j = 0; j < 8; ++j)
    {
        const int z = z0 + j;
        uint8_t l = 255, lb = 0;
        if (z < P.nz)
        {
            const int64_t i = base + z;
            if (kMode == 2)
            {
                l = lvl[i];
                lb = label[i] == 1;
            }
            else if (occ[i])
            {
                lb = 1;
                if (kMode == 0)
                {
                    const uint64_t m = lam2[i];
                    for (int q = 0; q < nlev; ++q)
                        if (m < lt.T[q])
                        {
                            l = (uint8_t)q;
                            break;
                        }
                }
                else
                {
                    const float m = meas[i];
                    for (int q = 0; q < nlev; ++q)
                        if (m < lf.L[q])
                        {
                            l = (uint8_t)q;
                            break;
                        }
                }
            }
        }
        lv8[j] = l;
        lb8[j] = lb;
        am |= (uint32_t)lb << j;
        em |= (uint32_t)(l <= k) << j;
        nm |= (uint32_t)(l == k) << j;
    }
    if (kMode != 2)
    {
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (z0 + j < P.nz)
            {
                lvl[base + z0 + j] = lv8[j];
                label[base + z0 + j] = lb8[j];
            }
    }
    const int sh = 8 * (lane & 7);
    uint64_t a = (uint64_t)am << sh, e = (uint64_t)em << sh,
             nw = (uint64_t)nm << sh;
#pragma unroll
    for (int d = 1; d <= 4; d <<= 1)
    {
        a |= __shfl_xor(a, d);
        e |= __shfl_xor(e, d);
        nw |= __shfl_xor(nw, d);
    }
    if ((lane & 7) == 0 && z0 < P.nz)
    {
        const int64_t o = rowi * P.nwz + (z0 >> 6);
        E[o] = e;
        const uint8_t f = (nw & a) ? 1 : 0;
        D0[o] = f;
        D1[o] = f;
        if (kMode != 2)
            A[o] = a;
    }
}

// bits z-1, z, z+1 of a row (prev/next = the neighbouring words' edge bits)
__device__ __forceinline__ uint32_t row3(uint64_t W, uint32_t prev63,
                                         uint32_t next0, int b)
{
    if (b == 0)
        return prev63 | (uint32_t)((W & 3ull) << 1);
    if (b == 63)
        return (uint32_t)(W >> 62) | (next0 << 2);
    return (uint32_t)((W >> (b - 1)) & 7ull);
}

// 26-neighbourhood code of bit b (bit order dx, dy, dz with dz fastest,
// centre skipped -- the order of the simple-point table)
__device__ __forceinline__ uint32_t nb_code(const uint64_t (&R)[3][3],
                                            const uint32_t (&pv)[3][3],
                                            const uint32_t (&nx0)[3][3], int b)
{
    uint32_t code = 0;
    int pos = 0;
#pragma unroll
    for (int dx = 0; dx < 3; ++dx)
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
        {
            const uint32_t r3 = row3(R[dx][dy], pv[dx][dy], nx0[dx][dy], b);
            if (dx == 1 && dy == 1)
            {
                code |= ((r3 & 1u) | ((r3 >> 1) & 2u)) << pos;
                pos += 2;
            }
            else
            {
                code |= r3 << pos;
                pos += 3;
            }
        }
    return code;
}

// Counters are spread over kSlots cache lines (one atomic per workgroup into
// slot blockIdx % kSlots): same-address device-scope atomics from every
// wavefront serialise across the XCDs.
constexpr int kSlots = 64;
constexpr int kSlotWords = 8 + VC_MAX_LEVELS; // [0] tests, [8 + k] removed

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* red)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        v += __shfl_xor(v, d);
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        red[wv] = v;
    __syncthreads();
    uint32_t tot = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i)
        tot += red[i];
    return tot;
}

// One subfield pass, dense over the words of the subfield's rows (thread =
// word; grid y = row x/2, grid x covers (y/2, w) with w padded to 2^lgw).
// A word is visited only when its z-parity flag is set: the flag of z-parity
// p (D0 / D1) is raised whenever a voxel of the 26-neighbourhood of one of
// its parity-p voxels is removed.  All flag traffic is plain byte stores: in
// pass (sx, sy, sz) the flags of parity sz are cleared only on rows of parity
// (sx, sy), and raised only on other rows.
__global__ void __launch_bounds__(kBlock)
    thin_pass_dense(uint64_t* A, const uint64_t* __restrict__ E,
                    uint8_t* __restrict__ label, uint8_t* D0, uint8_t* D1,
                    const uint32_t* __restrict__ lut, Plane P, int lgw, int s,
                    int k, int sweep, uint32_t* __restrict__ sweep_any,
                    unsigned long long* __restrict__ ctr)
{
    // a level ends after a sweep that removed nothing; later passes are no-ops
    if (sweep > 0 && sweep_any[sweep - 1] == 0)
        return;
    __shared__ uint32_t red[kBlock / 64];
    const int sx = s & 1, sy = (s >> 1) & 1, sz = (s >> 2) & 1;
    const int nyh = (P.ny - sy + 1) >> 1;
    const uint32_t tt = blockIdx.x * blockDim.x + threadIdx.x;
    const int w = (int)(tt & ((1u << lgw) - 1u));
    const int yh = (int)(tt >> lgw);
    const int x = 2 * (int)blockIdx.y + sx;
    const int y = 2 * yh + sy;
    uint8_t* Fs = sz ? D1 : D0; // this pass's parity
    uint8_t* Fo = sz ? D0 : D1; // the other parity
    uint64_t removed = 0;
    uint32_t ntest = 0;
    if (yh < nyh && w < P.nwz)
    {
        const int64_t o = widx(P, x, y, w);
        uint64_t C = 0, cand = 0;
        if (Fs[o])
        {
            Fs[o] = 0;
            C = A[o];
            const uint64_t zsub =
                sz ? 0xAAAAAAAAAAAAAAAAull : 0x5555555555555555ull;
            cand = C & E[o] & zsub;
        }
        if (cand)
        {
            uint64_t R[3][3];
            uint32_t pv[3][3], nx0[3][3];
            const bool need_prev = (cand & 1ull) && w > 0;
            const bool need_next = (cand >> 63) && w + 1 < P.nwz;
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
                for (int dy = -1; dy <= 1; ++dy)
                {
                    const bool in = (unsigned)(x + dx) < (unsigned)P.nx &&
                                    (unsigned)(y + dy) < (unsigned)P.ny;
                    const int64_t q = widx(P, x + dx, y + dy, w);
                    R[dx + 1][dy + 1] =
                        (dx == 0 && dy == 0) ? C : (in ? A[q] : 0ull);
                    pv[dx + 1][dy + 1] =
                        (in && need_prev) ? (uint32_t)(A[q - 1] >> 63) : 0u;
                    nx0[dx + 1][dy + 1] =
                        (in && need_next) ? (uint32_t)(A[q + 1] & 1ull) : 0u;
                }
            // border filter: a voxel with all six face neighbours alive is
            // never simple
            const uint64_t zlo = (C << 1) | (uint64_t)pv[1][1];
            const uint64_t zhi = (C >> 1) | ((uint64_t)nx0[1][1] << 63);
            cand &= ~(R[0][1] & R[2][1] & R[1][0] & R[1][2] & zlo & zhi);
            // four independent table lookups in flight per round
            while (cand)
            {
                int bb[4];
                uint32_t cc[4];
                int nb = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                {
                    bb[j] = 0;
                    cc[j] = 0;
                    if (cand)
                    {
                        bb[j] = __ffsll((unsigned long long)cand) - 1;
                        cand &= cand - 1;
                        cc[j] = nb_code(R, pv, nx0, bb[j]);
                        nb = j + 1;
                    }
                }
                uint32_t lw[4];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    lw[j] = j < nb ? lut[cc[j] >> 5] : 0u;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if ((lw[j] >> (cc[j] & 31)) & 1u)
                        removed |= 1ull << bb[j];
                ntest += nb;
            }
            if (removed)
            {
                A[o] = C & ~removed;
                uint8_t* row = label + ((int64_t)x * P.ny + y) * P.nz + 64 * w;
                uint64_t m = removed;
                while (m)
                {
                    const int b = __ffsll((unsigned long long)m) - 1;
                    m &= m - 1;
                    row[b] = (uint8_t)(2 + k);
                }
                // raise the flags of every word holding a neighbour
                const bool lo = (removed & 1ull) && w > 0;
                const bool hi = (removed >> 63) && w + 1 < P.nwz;
#pragma unroll
                for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
                    for (int dy = -1; dy <= 1; ++dy)
                    {
                        if ((unsigned)(x + dx) >= (unsigned)P.nx ||
                            (unsigned)(y + dy) >= (unsigned)P.ny)
                            continue;
                        const int64_t q = widx(P, x + dx, y + dy, w);
                        if (dx == 0 && dy == 0)
                        {
                            Fo[q] = 1;
                            if (lo)
                                Fo[q - 1] = 1;
                            if (hi)
                                Fo[q + 1] = 1;
                        }
                        else
                        {
                            D0[q] = 1;
                            D1[q] = 1;
                            if (lo)
                            {
                                D0[q - 1] = 1;
                                D1[q - 1] = 1;
                            }
                            if (hi)
                            {
                                D0[q + 1] = 1;
                                D1[q + 1] = 1;
                            }
                        }
                    }
            }
        }
    }
    const uint32_t nrem = block_sum((uint32_t)__popcll(removed), red);
    __syncthreads();
    const uint32_t ntst = block_sum(ntest, red);
    if (threadIdx.x == 0)
    {
        unsigned long long* c =
            ctr + (size_t)((blockIdx.x + blockIdx.y * gridDim.x) % kSlots) *
                      kSlotWords;
        if (nrem)
        {
            sweep_any[sweep] = 1u;
            atomicAdd(&c[8 + k], (unsigned long long)nrem);
        }
        if (ntst)
            atomicAdd(&c[0], (unsigned long long)ntst);
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
    Plane P{v->nx, v->ny, v->nz, (v->nz + 63) / 64};
    int lgw = 0;
    while ((1 << lgw) < P.nwz)
        lgw++;
    const size_t nwords = (size_t)v->nx * v->ny * P.nwz;
    int rc;
    if (!v->label && (rc = vc_alloc(ctx, (void**)&v->label, N)) != VC_OK)
        return rc;
    if (!v->lvl && (rc = vc_alloc(ctx, (void**)&v->lvl, N)) != VC_OK)
        return rc;
    if (!v->alive &&
        (rc = vc_alloc(ctx, (void**)&v->alive, 8 * nwords)) != VC_OK)
        return rc;
    if (!v->elig && (rc = vc_alloc(ctx, (void**)&v->elig, 8 * nwords)) != VC_OK)
        return rc;
    if (!v->dflag && (rc = vc_alloc(ctx, (void**)&v->dflag, 2 * nwords)) != VC_OK)
        return rc;
    const int kMaxSweeps = 1 << 16;
    if (!v->sweep_any &&
        (rc = vc_alloc(ctx, (void**)&v->sweep_any, 4 * (size_t)kMaxSweeps)) !=
            VC_OK)
        return rc;
    const size_t ctr_bytes = 8 * (size_t)kSlots * kSlotWords;
    if (!v->tstats &&
        (rc = vc_alloc(ctx, (void**)&v->tstats, ctr_bytes)) != VC_OK)
        return rc;
    if (!v->hcnt)
        VC_HIP(ctx, hipHostMalloc((void**)&v->hcnt, 4 * 64,
                                  hipHostMallocDefault));
    uint32_t* hsw = v->hcnt;

    hipStream_t S = ctx->stream;
    hipError_t e = hipMemsetAsync(v->tstats, 0, ctr_bytes, S);
    int64_t launches = 0;
    LevelsU64 lt{};
    LevelsF32 lf{};
    for (int k = 0; k < n; ++k)
    {
        lt.T[k] = (uint64_t)vc_level_threshold(levels[k]);
        lf.L[k] = levels[k];
    }
    const int nchunk = (v->nz + 511) / 512;
    const int64_t init_threads = (int64_t)v->nx * v->ny * nchunk * 64;
    const unsigned ginit = (unsigned)((init_threads + kBlock - 1) / kBlock);
    int64_t passes_total = 0;
    for (int k = 0; k < n && e == hipSuccess; ++k)
    {
        if (k > 0)
            thin_level_init<2><<<ginit, kBlock, 0, S>>>(
                v->occ, v->lam2, v->meas_f, v->lvl, v->label, v->alive, v->elig,
                v->dflag, v->dflag + nwords, P, nchunk, lt, lf, n, k);
        else if (v->measure_kind == 2)
            thin_level_init<1><<<ginit, kBlock, 0, S>>>(
                v->occ, v->lam2, v->meas_f, v->lvl, v->label, v->alive, v->elig,
                v->dflag, v->dflag + nwords, P, nchunk, lt, lf, n, k);
        else
            thin_level_init<0><<<ginit, kBlock, 0, S>>>(
                v->occ, v->lam2, v->meas_f, v->lvl, v->label, v->alive, v->elig,
                v->dflag, v->dflag + nwords, P, nchunk, lt, lf, n, k);
        launches++;
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemsetAsync(v->sweep_any, 0, 4 * (size_t)kMaxSweeps, S);
        // launch sweeps in batches; a sweep after the converged one is a
        // no-op (every pass exits at its first instruction)
        int sweep = 0;
        int batch = 4;
        while (e == hipSuccess)
        {
            if (sweep + batch > kMaxSweeps)
                return vc_fail(ctx, VC_ERR_GUARD,
                               "thinning did not converge in %d sweeps", sweep);
            for (int j = 0; j < batch && e == hipSuccess; ++j, ++sweep)
                for (int s = 0; s < 8 && e == hipSuccess; ++s)
                {
                    const int sx = s & 1, sy = (s >> 1) & 1;
                    const int nxh = (v->nx - sx + 1) >> 1;
                    const int nyh = (v->ny - sy + 1) >> 1;
                    if (nxh == 0 || nyh == 0)
                        continue;
                    const int64_t per_row = (int64_t)nyh << lgw;
                    const dim3 g((unsigned)((per_row + kBlock - 1) / kBlock),
                                 (unsigned)nxh);
                    thin_pass_dense<<<g, kBlock, 0, S>>>(
                        v->alive, v->elig, v->label, v->dflag,
                        v->dflag + nwords, ctx->d_lut, P, lgw, s, k, sweep,
                        v->sweep_any, (unsigned long long*)v->tstats);
                    launches++;
                    e = hipGetLastError();
                }
            if (e != hipSuccess)
                break;
            e = hipMemcpyAsync(hsw, v->sweep_any + sweep - batch,
                               4 * (size_t)batch, hipMemcpyDeviceToHost, S);
            if (e == hipSuccess)
                e = hipStreamSynchronize(S);
            if (e != hipSuccess)
                break;
            int conv = -1;
            for (int j = 0; j < batch; ++j)
                if (hsw[j] == 0)
                {
                    conv = sweep - batch + j;
                    break;
                }
            if (conv >= 0)
            {
                passes_total += 8LL * (conv + 1);
                break;
            }
            batch = batch < 16 ? batch * 2 : 16;
        }
    }
    if (e == hipSuccess && st)
    {
        std::vector<unsigned long long> hs((size_t)kSlots * kSlotWords);
        e = hipMemcpyAsync(hs.data(), v->tstats, ctr_bytes,
                           hipMemcpyDeviceToHost, S);
        if (e == hipSuccess)
            e = hipStreamSynchronize(S);
        if (e == hipSuccess)
        {
            std::memset(st, 0, sizeof(*st));
            for (int sl = 0; sl < kSlots; ++sl)
            {
                const unsigned long long* c = &hs[(size_t)sl * kSlotWords];
                st->tests += (int64_t)c[0];
                for (int k = 0; k < n; ++k)
                    st->removed[k] += (int64_t)c[8 + k];
            }
            st->passes = passes_total;
            st->launches = launches;
        }
    }
    if (e != hipSuccess)
        return vc_fail(ctx, VC_ERR_HIP, "thinning: %s", hipGetErrorString(e));
    v->thin_done = true;
    return VC_OK;
}
