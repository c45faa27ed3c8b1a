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

// one wavefront packs one word: lane l <-> z = 64w + l.  Words holding an
// alive voxel that becomes eligible at level k are flagged for both z
// parities (every other eligible voxel was found non-simple when the previous
// level converged and its neighbourhood has not changed since).
__global__ void __launch_bounds__(kBlock)
    pack_planes(const uint8_t* __restrict__ label, const uint8_t* __restrict__ lvl,
                uint64_t* __restrict__ A, uint64_t* __restrict__ E,
                uint8_t* __restrict__ D0, uint8_t* __restrict__ D1, Plane P,
                int k)
{
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int x = blockIdx.z;
    const bool in = z < P.nz;
    const int64_t i = ((int64_t)x * P.ny + y) * P.nz + z;
    const uint8_t l = in ? lvl[i] : (uint8_t)255;
    const uint64_t e = __ballot(l <= k);
    const uint64_t enew = __ballot(l == k);
    const uint64_t a = __ballot(in && label[i] == 1);
    // lane 0 sits at z = 64w: the word exists iff 64w < nz
    if ((threadIdx.x & 63) == 0 && z < P.nz)
    {
        const int64_t o = widx(P, x, y, z >> 6);
        E[o] = e;
        A[o] = a;
        const uint8_t f = (enew & a) ? 1 : 0;
        D0[o] = f;
        D1[o] = f;
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

// One subfield pass, dense over the words of the subfield's rows (thread =
// word).  A word is visited only when its z-parity flag is set: the flag of
// z-parity p (D0 / D1) is raised whenever a voxel of the 26-neighbourhood of
// one of its parity-p voxels is removed.  All flag traffic is plain byte
// stores: in pass (sx, sy, sz) the flags of parity sz are cleared only on
// rows of parity (sx, sy), and raised only on the other rows.
__global__ void __launch_bounds__(kBlock)
    thin_pass_dense(uint64_t* A, const uint64_t* __restrict__ E,
                    uint8_t* __restrict__ label, uint8_t* D0, uint8_t* D1,
                    const uint32_t* __restrict__ lut, Plane P, int s, int k,
                    int sweep, unsigned long long* __restrict__ sweep_removed,
                    unsigned long long* __restrict__ stats)
{
    // a level ends after a sweep that removed nothing; later passes are no-ops
    if (sweep > 0 && sweep_removed[sweep - 1] == 0)
        return;
    const int sx = s & 1, sy = (s >> 1) & 1, sz = (s >> 2) & 1;
    const int nxh = (P.nx - sx + 1) >> 1, nyh = (P.ny - sy + 1) >> 1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)nxh * nyh * P.nwz;
    uint8_t* Fs = sz ? D1 : D0; // this pass's parity
    uint8_t* Fo = sz ? D0 : D1; // the other parity
    uint64_t removed = 0;
    uint32_t ntest = 0;
    if (t < total)
    {
        const int w = (int)(t % P.nwz);
        const int64_t r = t / P.nwz;
        const int y = 2 * (int)(r % nyh) + sy;
        const int x = 2 * (int)(r / nyh) + sx;
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
                for (int q = 0; q < 4; ++q)
                {
                    bb[q] = 0;
                    cc[q] = 0;
                    if (cand)
                    {
                        bb[q] = __ffsll((unsigned long long)cand) - 1;
                        cand &= cand - 1;
                        cc[q] = nb_code(R, pv, nx0, bb[q]);
                        nb = q + 1;
                    }
                }
                uint32_t lw[4];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    lw[q] = q < nb ? lut[cc[q] >> 5] : 0u;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if ((lw[q] >> (cc[q] & 31)) & 1u)
                        removed |= 1ull << bb[q];
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
    // wavefront reductions of the counters
    uint32_t nrem = (uint32_t)__popcll(removed);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
    {
        nrem += __shfl_xor(nrem, o);
        ntest += __shfl_xor(ntest, o);
    }
    if ((threadIdx.x & 63) == 0)
    {
        if (nrem)
        {
            atomicAdd(&sweep_removed[sweep], (unsigned long long)nrem);
            atomicAdd(&stats[3 + k], (unsigned long long)nrem);
        }
        if (ntest)
            atomicAdd(&stats[0], (unsigned long long)ntest);
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
    if (!v->sweep_removed &&
        (rc = vc_alloc(ctx, (void**)&v->sweep_removed,
                       8 * (size_t)kMaxSweeps)) != VC_OK)
        return rc;
    if (!v->tstats &&
        (rc = vc_alloc(ctx, (void**)&v->tstats, 8 * (3 + VC_MAX_LEVELS))) !=
            VC_OK)
        return rc;
    if (!v->hcnt)
        VC_HIP(ctx, hipHostMalloc((void**)&v->hcnt, 8 * 64,
                                  hipHostMallocDefault));
    unsigned long long* hsw = (unsigned long long*)v->hcnt;
    unsigned long long* swr = (unsigned long long*)v->sweep_removed;

    hipStream_t S = ctx->stream;
    hipError_t e = hipMemsetAsync(v->tstats, 0, 8 * (3 + VC_MAX_LEVELS), S);
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
    const dim3 gpack((unsigned)((v->nz + kBlock - 1) / kBlock), (unsigned)v->ny,
                     (unsigned)v->nx);
    int64_t passes_total = 0;
    for (int k = 0; k < n && e == hipSuccess; ++k)
    {
        pack_planes<<<gpack, kBlock, 0, S>>>(v->label, v->lvl, v->alive,
                                              v->elig, v->dflag, v->dflag + nwords,
                                              P, k);
        launches++;
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemsetAsync(swr, 0, 8 * (size_t)kMaxSweeps, S);
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
                    const int64_t words = (int64_t)((v->nx - sx + 1) >> 1) *
                                          ((v->ny - sy + 1) >> 1) * P.nwz;
                    const unsigned g = (unsigned)((words + kBlock - 1) / kBlock);
                    if (g == 0)
                        continue;
                    thin_pass_dense<<<g, kBlock, 0, S>>>(
                        v->alive, v->elig, v->label, v->dflag,
                        v->dflag + nwords, ctx->d_lut, P, s, k, sweep, swr,
                        (unsigned long long*)v->tstats);
                    launches++;
                    e = hipGetLastError();
                }
            if (e != hipSuccess)
                break;
            e = hipMemcpyAsync(hsw, swr + sweep - batch, 8 * (size_t)batch,
                               hipMemcpyDeviceToHost, S);
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
        uint64_t hs[3 + VC_MAX_LEVELS];
        e = hipMemcpyAsync(hs, v->tstats, sizeof(hs), hipMemcpyDeviceToHost, S);
        if (e == hipSuccess)
            e = hipStreamSynchronize(S);
        if (e == hipSuccess)
        {
            std::memset(st, 0, sizeof(*st));
            st->tests = (int64_t)hs[0];
            st->passes = passes_total;
            st->launches = launches;
            for (int k = 0; k < n; ++k)
                st->removed[k] = (int64_t)hs[3 + k];
        }
    }
    if (e != hipSuccess)
        return vc_fail(ctx, VC_ERR_HIP, "thinning: %s", hipGetErrorString(e));
    v->thin_done = true;
    return VC_OK;
}
