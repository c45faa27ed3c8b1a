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
// words whose neighbourhood did not change since their last visit.  A volume's
// level ends after a sweep that removed nothing in it (its items skip later
// passes); a batch's level ends when that holds for every volume.  Sweeps are
// launched in segments (optionally one captured graph per segment length);
// the host checks the per-sweep "removed anything" flags once per segment.
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

// Level set-up.  An item is 4 wavefronts; a wavefront covers 512 consecutive
// z of one row (8 words): lane l handles z = z0 + 8l .. z0 + 8l + 7 and lanes
// 8j .. 8j+7 assemble word j.  Words holding an alive voxel that becomes
// eligible at level k are flagged for both z parities (every other eligible
// voxel was found non-simple when the previous level converged and its
// neighbourhood has not changed since).  Level 0 derives each voxel's first
// eligible level from its volume's thresholds (exact: lam2 < T[q]; float
// measure: m < L[q]); later levels read it back.  The launch also resets the
// level's sweep control (sweep counter, level, "removed anything" rings).
constexpr int kLiWaves = 4;

template <bool kB>
__global__ void __launch_bounds__(kBlock)
    thin_level_init(vc_work W, const uint64_t* __restrict__ Tq,
                    const float* __restrict__ Lq, int nlev, int k, int nvol,
                    uint32_t* __restrict__ ctl, uint32_t* __restrict__ gany,
                    uint32_t* __restrict__ vany)
{
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    {
        const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
        const uint32_t gs = gridDim.x * blockDim.x;
        if (g == 0)
        {
            ctl[0] = 0;
            ctl[1] = (uint32_t)k;
        }
        for (uint32_t i = g; i < (uint32_t)VC_RING; i += gs)
            gany[i] = 0;
        for (uint32_t i = g; i < 4u * (uint32_t)nvol; i += gs)
            vany[i] = 0;
    }
    for (uint32_t b = blockIdx.x; b < W.total; b += gridDim.x)
    {
        int vi;
        uint32_t lt;
        vc_item<kB>(W, b, vi, lt);
        const vc_vdesc d = vc_desc<kB>(W, vi);
        const int nz = d.nz;
        const int nchunk = (nz + 511) >> 9;
        const int64_t wave = (int64_t)lt * kLiWaves + wv;
        const int64_t rowi = wave / nchunk;
        if (rowi >= (int64_t)d.nx * d.ny) // whole wavefront
            continue;
        const int z0 = (int)(wave % nchunk) * 512 + 8 * lane;
        const int64_t base = rowi * nz;
        const int mode = k > 0 ? 2 : (d.mkind == 2 ? 1 : 0);
        // rows start 8-aligned when nz is a multiple of 8: one 8-byte access
        // per lane instead of eight byte accesses (z0 + 8 <= nz then)
        const bool vec = (nz & 7) == 0 && z0 < nz;
        uint32_t am = 0, em = 0, nm = 0;
        uint8_t lv8[8], lb8[8];
        if (mode == 2)
        {
            // alive voxels from the maintained plane; levels read back
            if (z0 < nz)
                am = (uint32_t)(d.alive[rowi * d.nwz + (z0 >> 6)] >> (z0 & 63)) &
                     0xFFu;
            if (vec)
            {
                const uint64_t v =
                    *reinterpret_cast<const uint64_t*>(d.lvl + base + z0);
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    lv8[j] = (uint8_t)(v >> (8 * j));
            }
            else
            {
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    lv8[j] = z0 + j < nz ? d.lvl[base + z0 + j] : (uint8_t)255;
            }
#pragma unroll
            for (int j = 0; j < 8; ++j)
                lb8[j] = (uint8_t)((am >> j) & 1u);
        }
        else
        {
            // measure of the inside voxels, then the first level whose
            // threshold exceeds it (thresholds ascending; q descending so the
            // smallest such q is kept); thresholds are uniform scalar loads
            uint64_t mi[8];
            float mf[8];
            if (vec)
            {
                const uint64_t ov =
                    *reinterpret_cast<const uint64_t*>(d.occ + base + z0);
                uint32_t m32[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                // 8 background voxels need no measure: skipping their loads
                // leaves whole sectors of the measure unread inside the
                // background (70 % of the 512^3 headline volume)
                if (mode == 0 && ov)
                {
                    const uint4 u0 = *reinterpret_cast<const uint4*>(d.lam2 + base + z0);
                    const uint4 u1 =
                        *reinterpret_cast<const uint4*>(d.lam2 + base + z0 + 4);
                    m32[0] = u0.x; m32[1] = u0.y; m32[2] = u0.z; m32[3] = u0.w;
                    m32[4] = u1.x; m32[5] = u1.y; m32[6] = u1.z; m32[7] = u1.w;
                }
                else if (mode != 0 && ov)
                {
                    const float4 f0 = *reinterpret_cast<const float4*>(d.meas + base + z0);
                    const float4 f1 =
                        *reinterpret_cast<const float4*>(d.meas + base + z0 + 4);
                    mf[0] = f0.x; mf[1] = f0.y; mf[2] = f0.z; mf[3] = f0.w;
                    mf[4] = f1.x; mf[5] = f1.y; mf[6] = f1.z; mf[7] = f1.w;
                }
#pragma unroll
                for (int j = 0; j < 8; ++j)
                {
                    lb8[j] = ((ov >> (8 * j)) & 0xFFu) != 0;
                    lv8[j] = 255;
                    mi[j] = lb8[j] ? m32[j] : 0;
                    if (mode == 0 || !lb8[j])
                        mf[j] = 0.0f;
                }
            }
            else
            {
#pragma unroll
                for (int j = 0; j < 8; ++j)
                {
                    const int64_t i = base + z0 + j;
                    const bool in = z0 + j < nz && d.occ[i] != 0;
                    lb8[j] = in;
                    lv8[j] = 255;
                    mi[j] = 0;
                    mf[j] = 0.0f;
                    if (in)
                    {
                        if (mode == 0)
                            mi[j] = d.lam2[i];
                        else
                            mf[j] = d.meas[i];
                    }
                }
            }
            typedef const __attribute__((address_space(4))) uint64_t* cu64;
            typedef const __attribute__((address_space(4))) float* cf32;
            const cu64 T = (cu64)(Tq + (size_t)vi * VC_MAX_LEVELS);
            const cf32 Lf = (cf32)(Lq + (size_t)vi * VC_MAX_LEVELS);
            for (int q = nlev - 1; q >= 0; --q)
            {
                if (mode == 0)
                {
                    const uint64_t t = T[q];
#pragma unroll
                    for (int j = 0; j < 8; ++j)
                        if (lb8[j] && mi[j] < t)
                            lv8[j] = (uint8_t)q;
                }
                else
                {
                    const float t = Lf[q];
#pragma unroll
                    for (int j = 0; j < 8; ++j)
                        if (lb8[j] && mf[j] < t)
                            lv8[j] = (uint8_t)q;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < 8; ++j)
        {
            am |= (uint32_t)lb8[j] << j;
            em |= (uint32_t)(lv8[j] <= k) << j;
            nm |= (uint32_t)(lv8[j] == k) << j;
        }
        if (mode != 2)
        {
            if (vec)
            {
                uint64_t lvw = 0, lbw = 0;
#pragma unroll
                for (int j = 0; j < 8; ++j)
                {
                    lvw |= (uint64_t)lv8[j] << (8 * j);
                    lbw |= (uint64_t)lb8[j] << (8 * j);
                }
                *reinterpret_cast<uint64_t*>(d.lvl + base + z0) = lvw;
                *reinterpret_cast<uint64_t*>(d.label + base + z0) = lbw;
            }
            else
            {
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (z0 + j < nz)
                    {
                        d.lvl[base + z0 + j] = lv8[j];
                        d.label[base + z0 + j] = lb8[j];
                    }
            }
        }
        const int sh = 8 * (lane & 7);
        uint64_t a = (uint64_t)am << sh, e = (uint64_t)em << sh,
                 nw = (uint64_t)nm << sh;
#pragma unroll
        for (int dd = 1; dd <= 4; dd <<= 1)
        {
            a |= __shfl_xor(a, dd);
            e |= __shfl_xor(e, dd);
            nw |= __shfl_xor(nw, dd);
        }
        if ((lane & 7) == 0 && z0 < nz)
        {
            const int64_t o = rowi * d.nwz + (z0 >> 6);
            d.elig[o] = e;
            const uint8_t f = (nw & a) ? 1 : 0;
            d.d0[o] = f;
            d.d1[o] = f;
            if (mode != 2)
                d.alive[o] = a;
        }
    }
}

// Level set-up for k > 0, one thread per plane word: elig = (lvl <= k) of
// its 64 voxels, and both revisit flags = an alive voxel becomes eligible
// now (lvl == k).  The 64 level bytes are read as four 16-byte loads when the
// rows are 16-byte aligned (nz a multiple of 16).  Also resets the level's
// sweep control.
template <bool kB>
__global__ void __launch_bounds__(kBlock)
    thin_level_next(vc_work W, int k, int nvol, uint32_t* __restrict__ ctl,
                    uint32_t* __restrict__ gany, uint32_t* __restrict__ vany)
{
    {
        const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
        const uint32_t gs = gridDim.x * blockDim.x;
        if (g == 0)
        {
            ctl[0] = 0;
            ctl[1] = (uint32_t)k;
        }
        for (uint32_t i = g; i < (uint32_t)VC_RING; i += gs)
            gany[i] = 0;
        for (uint32_t i = g; i < 4u * (uint32_t)nvol; i += gs)
            vany[i] = 0;
    }
    for (uint32_t b = blockIdx.x; b < W.total; b += gridDim.x)
    {
        int vi;
        uint32_t lt;
        vc_item<kB>(W, b, vi, lt);
        const vc_vdesc d = vc_desc<kB>(W, vi);
        const int64_t o = (int64_t)lt * kBlock + threadIdx.x; // plane word
        const int64_t words = (int64_t)d.nx * d.ny * d.nwz;
        if (o >= words)
            continue;
        const int64_t rowi = o / d.nwz;
        const int z0 = (int)(o - rowi * d.nwz) * 64;
        const uint8_t* lv = d.lvl + rowi * d.nz + z0;
        uint64_t e = 0, nw = 0;
        // eligibility only matters for alive voxels (every use ands it with
        // the alive plane): a word without one reads no levels, which at
        // later levels, after most voxels were removed, skips most of them
        const uint64_t al = d.alive[o];
        if (!al)
        {
        }
        else if ((d.nz & 15) == 0 && z0 + 64 <= d.nz)
        {
#pragma unroll
            for (int c = 0; c < 4; ++c)
            {
                const uint4 q = reinterpret_cast<const uint4*>(lv)[c];
                const uint32_t wd[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                    {
                        const uint32_t l = (wd[u] >> (8 * j)) & 0xFFu;
                        const int bit = 16 * c + 4 * u + j;
                        e |= (uint64_t)(l <= (uint32_t)k) << bit;
                        nw |= (uint64_t)(l == (uint32_t)k) << bit;
                    }
            }
        }
        else
        {
            const int n = min(64, d.nz - z0);
            for (int j = 0; j < n; ++j)
            {
                const uint32_t l = lv[j];
                e |= (uint64_t)(l <= (uint32_t)k) << j;
                nw |= (uint64_t)(l == (uint32_t)k) << j;
            }
        }
        d.elig[o] = e;
        const uint8_t f = (nw & al) ? 1 : 0;
        d.d0[o] = f;
        d.d1[o] = f;
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

// 3x3 rows of alive words around word (x, y, w) and the edge bits of their z
// neighbours (prev: bit 63 of word w-1, next: bit 0 of word w+1, loaded only
// when a candidate sits at that end).  Returns the candidates that pass the
// border filter: a voxel with all six face neighbours alive is never simple.
struct Nbhd
{
    uint64_t R[3][3];
    uint32_t pv[3][3], nx0[3][3];
};

__device__ __forceinline__ uint64_t load_nbhd(const Plane& P, const uint64_t* A,
                                              uint64_t C, uint64_t cand, int x,
                                              int y, int w, Nbhd& N)
{
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
            N.R[dx + 1][dy + 1] = (dx == 0 && dy == 0) ? C : (in ? A[q] : 0ull);
            N.pv[dx + 1][dy + 1] =
                (in && need_prev) ? (uint32_t)(A[q - 1] >> 63) : 0u;
            N.nx0[dx + 1][dy + 1] =
                (in && need_next) ? (uint32_t)(A[q + 1] & 1ull) : 0u;
        }
    const uint64_t zlo = (C << 1) | (uint64_t)N.pv[1][1];
    const uint64_t zhi = (C >> 1) | ((uint64_t)N.nx0[1][1] << 63);
    return cand & ~(N.R[0][1] & N.R[2][1] & N.R[1][0] & N.R[1][2] & zlo & zhi);
}

// Removal of `removed` from word o (value C): the word, the labels, and the
// revisit flags of every word holding a neighbour.
//
// Revisit flags: the flag of z-parity p (D0 / D1) is raised whenever a voxel
// of the 26-neighbourhood of one of the word's parity-p voxels is removed.
// All flag traffic is plain byte stores: in pass (sx, sy, sz) the flags of
// parity sz are cleared only on rows of parity (sx, sy), and raised only on
// other rows.
__device__ __forceinline__ void apply_removal(const Plane& P, uint64_t* A,
                                              uint8_t* __restrict__ label,
                                              uint8_t* D0, uint8_t* D1,
                                              uint8_t* Fo, int x, int y, int w,
                                              int64_t o, uint64_t C,
                                              uint64_t removed, int k)
{
    A[o] = C & ~removed;
    uint8_t* row = label + ((int64_t)x * P.ny + y) * P.nz + 64 * w;
    uint64_t m = removed;
    while (m)
    {
        const int bt = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        row[bt] = (uint8_t)(2 + k);
    }
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

// Visit of one word by one thread (sparse pass; its flag was set and has
// been cleared): candidates of the pass's z parity (alive, eligible, a
// background face neighbour) go through the simple-point table, four lookups
// in flight per round.  Returns the removed bits; *ntest counts the lookups.
__device__ __forceinline__ uint64_t visit_word(
    const Plane& P, uint64_t* A, const uint64_t* __restrict__ E,
    uint8_t* __restrict__ label, uint8_t* D0, uint8_t* D1, uint8_t* Fo,
    const uint32_t* __restrict__ lut, int x, int y, int w, int64_t o,
    uint64_t zsub, int k, uint32_t* ntest)
{
    const uint64_t C = A[o];
    uint64_t cand = C & E[o] & zsub;
    if (!cand)
        return 0;
    Nbhd N;
    cand = load_nbhd(P, A, C, cand, x, y, w, N);
    uint64_t removed = 0;
    while (cand)
    {
        int bb[4];
        uint32_t cc[4];
        int nb = 0;
#pragma unroll
        for (int jj = 0; jj < 4; ++jj)
        {
            bb[jj] = 0;
            cc[jj] = 0;
            if (cand)
            {
                bb[jj] = __ffsll((unsigned long long)cand) - 1;
                cand &= cand - 1;
                cc[jj] = nb_code(N.R, N.pv, N.nx0, bb[jj]);
                nb = jj + 1;
            }
        }
        uint32_t lw[4];
#pragma unroll
        for (int jj = 0; jj < 4; ++jj)
            lw[jj] = jj < nb ? lut[cc[jj] >> 5] : 0u;
#pragma unroll
        for (int jj = 0; jj < 4; ++jj)
            if ((lw[jj] >> (cc[jj] & 31)) & 1u)
                removed |= 1ull << bb[jj];
        *ntest += nb;
    }
    if (removed)
        apply_removal(P, A, label, D0, D1, Fo, x, y, w, o, C, removed, k);
    return removed;
}

// exclusive prefix sum of v over the workgroup (kBlock threads); *total gets
// the sum.  `wsum` holds kBlock/64 entries; ends after a barrier, so the
// caller must not rewrite wsum before its next barrier.
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t* wsum,
                                               uint32_t* total)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1)
    {
        const uint32_t t = __shfl_up(x, d);
        if (lane >= d)
            x += t;
    }
    if (lane == 63)
        wsum[wv] = x;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kBlock / 64; ++i)
    {
        const uint32_t t = wsum[i];
        base += i < wv ? t : 0u;
        tot += t;
    }
    *total = tot;
    return base + x - v;
}

// class-relative word t of a pass -> (x, y, w); false past the last row
__device__ __forceinline__ bool pass_word(const Plane& P, int sx, int sy,
                                          uint32_t t, int& x, int& y, int& w)
{
    const int nxh = (P.nx - sx + 1) >> 1, nyh = (P.ny - sy + 1) >> 1;
    const uint32_t r = t / (uint32_t)P.nwz;
    w = (int)(t - r * (uint32_t)P.nwz);
    const uint32_t xh = r / (uint32_t)nyh;
    const int yh = (int)(r - xh * (uint32_t)nyh);
    x = 2 * (int)xh + sx;
    y = 2 * yh + sy;
    return (int)xh < nxh;
}

// Sweep control lives on the device so that a segment of sweeps can be
// replayed as a captured graph: sweep = ctl[0] + j, level = ctl[1].  A volume
// whose previous sweep removed nothing has converged at this level (its next
// sweep would visit no word) and is skipped.  The "removed anything" flags
// are rings: vany[4v + (sweep & 3)] per volume, gany[sweep % VC_RING] for the
// batch; the first pass of a sweep clears the entries of sweep + 1 (vany) and
// sweep + VC_RING/2 (gany), which no pass of this sweep reads or writes.
struct PassCtl
{
    uint32_t sweep;
    int k;
};

__device__ __forceinline__ PassCtl pass_ctl(const uint32_t* ctl, int j,
                                            int k_arg, int s, int nvol,
                                            uint32_t* gany, uint32_t* vany)
{
    // captured: sweep and level from the device; launched: the arguments
    PassCtl c;
    c.sweep = ctl ? ctl[0] + (uint32_t)j : (uint32_t)j;
    c.k = ctl ? (int)ctl[1] : k_arg;
    if (s == 0)
    {
        const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
        const uint32_t gs = gridDim.x * blockDim.x;
        for (uint32_t i = g; i < (uint32_t)nvol; i += gs)
            vany[4 * i + ((c.sweep + 1) & 3)] = 0;
        if (g == 0)
            gany[(c.sweep + VC_RING / 2) & (VC_RING - 1)] = 0;
    }
    return c;
}

// One subfield pass, dense: an item is 256 words of one volume (thread =
// word, words ordered (x/2, y/2, w) within the pass's row parity).  Each
// thread filters its word's candidates; the candidates of the whole item are
// then queued in LDS (26-neighbourhood code and bit, owner word) and every
// thread takes kQ/kBlock queue entries per round, so the table lookups of a
// word with many candidates are spread over the workgroup and issued
// together.  Hits are or-ed into the owner word's removal mask in LDS; each
// owner applies its own removals.
constexpr int kQ = 1024;

template <bool kB>
__global__ void __launch_bounds__(kBlock)
    thin_pass(vc_work W, const uint32_t* __restrict__ lut, int s, int j,
              int k_arg, int nvol, const uint32_t* __restrict__ ctl,
              uint32_t* gany, uint32_t* vany,
              unsigned long long* __restrict__ ctr)
{
    const PassCtl pc = pass_ctl(ctl, j, k_arg, s, nvol, gany, vany);
    const uint32_t sweep = pc.sweep;
    const int k = pc.k;
    __shared__ uint32_t red[kBlock / 64];
    __shared__ uint32_t qc[kQ];  // code << 6 | bit
    __shared__ uint8_t qw[kQ];   // owner thread
    __shared__ unsigned long long rmv[kBlock];
    const int sx = s & 1, sy = (s >> 1) & 1, sz = (s >> 2) & 1;
    const uint64_t zsub = sz ? 0xAAAAAAAAAAAAAAAAull : 0x5555555555555555ull;
    for (uint32_t b = blockIdx.x; b < W.total; b += gridDim.x)
    {
        int vi;
        uint32_t lt;
        vc_item<kB>(W, b, vi, lt);
        const vc_vdesc d = vc_desc<kB>(W, vi);
        const Plane P{d.nx, d.ny, d.nz, d.nwz};
        uint8_t* Fs = sz ? d.d1 : d.d0; // this pass's parity
        uint8_t* Fo = sz ? d.d0 : d.d1; // the other parity
        int x, y, w;
        const bool valid =
            pass_word(P, sx, sy, lt * kBlock + threadIdx.x, x, y, w);
        const int64_t o = widx(P, x, y, w);
        // the flag load is issued before the convergence check so that the
        // two latencies overlap
        const uint8_t fl = valid ? Fs[o] : (uint8_t)0;
        if (sweep > 0 && vany[4 * vi + ((sweep - 1) & 3)] == 0)
            continue; // converged at this level (block-uniform)
        uint64_t C = 0, cand = 0;
        Nbhd N;
        if (fl)
        {
            Fs[o] = 0;
            C = d.alive[o];
            cand = C & d.elig[o] & zsub;
            if (cand)
                cand = load_nbhd(P, d.alive, C, cand, x, y, w, N);
        }
        // visited words (flagged), each standing for the 4 x 64 voxels of
        // its (x/2, y/2, w) cell; items without candidates end here
        const uint32_t nvis = (uint32_t)__syncthreads_count(fl != 0);
        const bool anycand = __syncthreads_or(cand != 0);
        if (!anycand)
        {
            if (threadIdx.x == 0 && nvis)
                atomicAdd(&ctr[(size_t)(b % VC_CTR_SLOTS) * VC_CTR_WORDS + 1],
                          (unsigned long long)nvis * 256ull);
            continue;
        }
        rmv[threadIdx.x] = 0;
        uint32_t total;
        const uint32_t off =
            block_scan((uint32_t)__popcll(cand), red, &total);
        uint32_t idx = off;
        uint64_t rest = cand;
        for (uint32_t c0 = 0; c0 < total; c0 += kQ)
        {
            // queue this thread's candidates that fall in [c0, c0 + kQ)
            while (rest && idx < c0 + kQ)
            {
                const int bt = __ffsll((unsigned long long)rest) - 1;
                rest &= rest - 1;
                qc[idx - c0] = (nb_code(N.R, N.pv, N.nx0, bt) << 6) | (uint32_t)bt;
                qw[idx - c0] = (uint8_t)threadIdx.x;
                ++idx;
            }
            __syncthreads();
            const uint32_t n = min((uint32_t)kQ, total - c0);
            uint32_t qv[kQ / kBlock], lw[kQ / kBlock];
#pragma unroll
            for (int r = 0; r < kQ / kBlock; ++r)
            {
                const uint32_t e = threadIdx.x + r * kBlock;
                qv[r] = e < n ? qc[e] : 0u;
            }
#pragma unroll
            for (int r = 0; r < kQ / kBlock; ++r)
            {
                const uint32_t e = threadIdx.x + r * kBlock;
                lw[r] = e < n ? lut[qv[r] >> 11] : 0u;
            }
#pragma unroll
            for (int r = 0; r < kQ / kBlock; ++r)
            {
                const uint32_t e = threadIdx.x + r * kBlock;
                if (e < n && ((lw[r] >> ((qv[r] >> 6) & 31u)) & 1u))
                    atomicOr(&rmv[qw[e]], 1ull << (qv[r] & 63u));
            }
            __syncthreads();
        }
        const uint64_t removed = rmv[threadIdx.x];
        if (removed)
            apply_removal(P, d.alive, d.label, d.d0, d.d1, Fo, x, y, w, o, C,
                          removed, k);
        const uint32_t nrem = block_sum((uint32_t)__popcll(removed), red);
        if (threadIdx.x == 0)
        {
            // counters spread over VC_CTR_SLOTS cache lines: same-address
            // device-scope atomics from every workgroup serialise across XCDs
            unsigned long long* c =
                ctr + (size_t)(b % VC_CTR_SLOTS) * VC_CTR_WORDS;
            if (nrem)
            {
                gany[sweep & (VC_RING - 1)] = 1u;
                vany[4 * vi + (sweep & 3)] = 1u;
                atomicAdd(&c[8 + k], (unsigned long long)nrem);
            }
            atomicAdd(&c[0], (unsigned long long)total);
            atomicAdd(&c[1], (unsigned long long)nvis * 256ull);
        }
        __syncthreads(); // red, rmv are rewritten by the next item
    }
}

// One subfield pass, sparse (late sweeps, few flagged words): one wavefront
// per item of 256 words; each lane reads 4 flags, the flagged words are
// compacted in LDS and the wavefront visits only those.  A quarter of the
// dense pass's wavefronts, each with one round of flag loads.
template <bool kB>
__global__ void __launch_bounds__(64)
    thin_pass_sparse(vc_work W, const uint32_t* __restrict__ lut, int s, int j,
                     int k_arg, int nvol, const uint32_t* __restrict__ ctl,
                     uint32_t* gany, uint32_t* vany,
                     unsigned long long* __restrict__ ctr)
{
    const PassCtl pc = pass_ctl(ctl, j, k_arg, s, nvol, gany, vany);
    const uint32_t sweep = pc.sweep;
    const int k = pc.k;
    __shared__ uint32_t list[kBlock];
    const int lane = threadIdx.x;
    const int sx = s & 1, sy = (s >> 1) & 1, sz = (s >> 2) & 1;
    const uint64_t zsub = sz ? 0xAAAAAAAAAAAAAAAAull : 0x5555555555555555ull;
    const uint64_t below = (1ull << lane) - 1ull;
    for (uint32_t b = blockIdx.x; b < W.total; b += gridDim.x)
    {
        int vi;
        uint32_t lt;
        vc_item<kB>(W, b, vi, lt);
        const vc_vdesc d = vc_desc<kB>(W, vi);
        const Plane P{d.nx, d.ny, d.nz, d.nwz};
        uint8_t* Fs = sz ? d.d1 : d.d0;
        uint8_t* Fo = sz ? d.d0 : d.d1;
        uint8_t fl[4];
        int64_t oo[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            int x, y, w;
            const bool valid =
                pass_word(P, sx, sy, lt * kBlock + q * 64 + lane, x, y, w);
            oo[q] = widx(P, x, y, w);
            fl[q] = valid ? Fs[oo[q]] : (uint8_t)0;
        }
        if (sweep > 0 && vany[4 * vi + ((sweep - 1) & 3)] == 0)
            continue; // converged at this level (block-uniform)
        uint32_t n = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            const uint64_t m = __ballot(fl[q] != 0);
            if (fl[q])
            {
                list[n + __popcll(m & below)] = (uint32_t)(q * 64 + lane);
                Fs[oo[q]] = 0;
            }
            n += (uint32_t)__popcll(m);
        }
        __syncthreads();
        uint32_t nrem = 0, ntest = 0;
        for (uint32_t i = lane; i < n; i += 64)
        {
            int x, y, w;
            pass_word(P, sx, sy, lt * kBlock + list[i], x, y, w);
            const uint64_t rm =
                visit_word(P, d.alive, d.elig, d.label, d.d0, d.d1, Fo, lut, x,
                           y, w, widx(P, x, y, w), zsub, k, &ntest);
            nrem += (uint32_t)__popcll(rm);
        }
#pragma unroll
        for (int dd = 32; dd >= 1; dd >>= 1)
        {
            nrem += __shfl_xor(nrem, dd);
            ntest += __shfl_xor(ntest, dd);
        }
        if (lane == 0)
        {
            unsigned long long* c =
                ctr + (size_t)(b % VC_CTR_SLOTS) * VC_CTR_WORDS;
            if (nrem)
            {
                gany[sweep & (VC_RING - 1)] = 1u;
                vany[4 * vi + (sweep & 3)] = 1u;
                atomicAdd(&c[8 + k], (unsigned long long)nrem);
            }
            if (ntest)
                atomicAdd(&c[0], (unsigned long long)ntest);
            if (n)
                atomicAdd(&c[1], (unsigned long long)n * 256ull);
        }
        __syncthreads(); // list is rewritten by the next item
    }
}

// ------------------------------------------------------------ tile sweeps
// One launch per sweep.  A tile is 32 x 32 rows (x, y) of one word (64 z).
// The sweep's 8 subfield passes run in LDS on the tile's region: the tile
// plus a halo of 8 rows in x, 4 rows in y and 2 bits in z.  Why that halo:
// the decision of voxel u at pass p = sx + 2 sy + 4 sz reaches the interior
// only through a chain of neighbours decided at strictly later passes; along
// such a chain x parity can flip at every pass, y parity at 1->2, 3->4, 5->6
// and z parity at 3->4 only.  So u matters iff its distance from the interior
// is at most Fx(p) = 7 - p in x, Fy(p) = {3,3,2,2,1,1,0,0}[p] in y and
// Fz(p) = (p < 4) in z; exactly those voxels are evaluated, and they read
// neighbours at most one further out, all inside the region.  A neighbour
// that changes before the reading pass is itself within its own pass's
// range (its F is larger by one in every axis in which it differs), so every
// read sees the true state and the interior is exact.
//
// Rows are held as 68-bit windows (bit i = z 64w - 2 + i) in two parts, a
// 64-bit low word and the 4 high bits.
//
// Planes: sweep t reads alive (t even) or alive2 (t odd) and writes the other
// one.  A tile that is not visited has equal content in both planes (it did
// not change in the previous sweep, and the sweep before wrote it), so no
// copy is needed.  A level's first sweep visits every tile, so it writes
// every tile of alive2; when tile sweeps start later in a level (after dense
// passes) the entry kernel copies alive to alive2 instead.  After the first
// sweep a tile is visited iff a tile of its 3 x 3 x 3 block changed in the
// previous sweep (changes travel at most 8 voxels per sweep, less than a
// tile); after an entry kernel the tiles holding a revisit flag count as
// changed.  At the end of a level (a sweep that changed nothing) both planes
// are equal again.
constexpr int kTT = 32;                     // interior rows along x and y
constexpr int kHX = 8, kHY = 4, kHZ = 2;    // halo: rows x, rows y, bits z
constexpr int kRX = kTT + 2 * kHX, kRY = kTT + 2 * kHY, kTRows = kRX * kRY;
// neighbour masks: bit (dx+1)*9 + (dy+1)*3 + (dz+1) for the tile at offset
// (dx, dy, dz); bit 13 is the tile itself
constexpr uint32_t kAllNbrs = (1u << 27) - 1u;

// The tiles whose region (tile + halo) holds a voxel at tile-local
// (x, y, z): the tile itself, plus the lower neighbour along an axis when the
// voxel lies within that neighbour's halo reach (x < kHX, ...), the upper one
// when x >= kTT - kHX, and the products of those along several axes.
__device__ __forceinline__ uint32_t halo_nbrs(int x, int y, int z)
{
    const uint32_t xs = 2u | (x < kHX ? 1u : 0u) | (x >= kTT - kHX ? 4u : 0u);
    const uint32_t ys = 2u | (y < kHY ? 1u : 0u) | (y >= kTT - kHY ? 4u : 0u);
    const uint32_t zs = 2u | (z < kHZ ? 1u : 0u) | (z >= 64 - kHZ ? 4u : 0u);
    const uint32_t yz = ((ys & 1u) ? zs : 0u) | ((ys & 2u) ? zs << 3 : 0u) |
                        ((ys & 4u) ? zs << 6 : 0u);
    return ((xs & 1u) ? yz : 0u) | ((xs & 2u) ? yz << 9 : 0u) |
           ((xs & 4u) ? yz << 18 : 0u);
}
#ifndef VC_TILE_BLOCK
#define VC_TILE_BLOCK 512
#endif
constexpr int kTBlock = VC_TILE_BLOCK; // threads of a tile-sweep workgroup
#ifndef VC_TILE_WAVES
#define VC_TILE_WAVES 8
#endif
// waves per SIMD the tile sweep is compiled for (4 workgroups of 512 per CU,
// which its 37 KB of LDS allows: 64 VGPRs)
constexpr int kTWaves = VC_TILE_WAVES;
constexpr int kTQ = 1024;              // candidate queue entries per round
// rows of one parity evaluated per pass: at most 23 x 19 (pass 0)
constexpr int kTRowsPerThread = (23 * 19 + kTBlock - 1) / kTBlock;
// a row holds at most 34 candidates of one z parity (68-bit window), and the
// scan's ballot prefix takes per-thread counts below 64
static_assert(kTRowsPerThread * 34 < 64, "tile_scan counts fit 6 bits");
__device__ __forceinline__ int fy_of(int p) { return (7 - p) >> 1; }

// window of word (o, w): bits 62..63 of word w-1, word w, bits 0..1 of w+1
__device__ __forceinline__ void load_win(const uint64_t* __restrict__ P,
                                         int64_t o, int w, int nwz,
                                         uint64_t& lo, uint32_t& hi)
{
    const uint64_t c = P[o];
    const uint64_t p = w > 0 ? P[o - 1] : 0ull;
    const uint64_t n = w + 1 < nwz ? P[o + 1] : 0ull;
    lo = (p >> 62) | (c << 2);
    hi = (uint32_t)(c >> 62) | ((uint32_t)(n & 3ull) << 2);
}

// a region row: window bits 0..63 and 64..67 (in LDS the low words and the
// high nibbles live in separate arrays, which keeps the tile sweep at 37 KB)
struct TRow
{
    uint64_t lo;
    uint32_t hi, pad;
};

// bits j, j+1, j+2 of a row given as dwords d0 (bits 0..31), d1, d2
__device__ __forceinline__ uint32_t bits3(uint32_t d0, uint32_t d1, uint32_t d2,
                                          int sel, int sh)
{
    const uint32_t lo = sel == 0 ? d0 : (sel == 1 ? d1 : d2);
    const uint32_t hi = sel == 0 ? d1 : (sel == 1 ? d2 : 0u);
    return __builtin_amdgcn_alignbit(hi, lo, (uint32_t)sh) & 7u;
}

// 26-neighbourhood code of window bit i (1 <= i <= 66) from the 3 x 3 rows
// around it (R[dx][dy]), the order of nb_code
__device__ __forceinline__ uint32_t nb_code_rows(const TRow (&R)[3][3], int i)
{
    const int j = i - 1, sel = j >> 5, sh = j & 31;
    uint32_t code = 0;
    int pos = 0;
#pragma unroll
    for (int dx = 0; dx < 3; ++dx)
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
        {
            const TRow& t = R[dx][dy];
            const uint32_t r3 = bits3((uint32_t)t.lo, (uint32_t)(t.lo >> 32),
                                      t.hi, sel, sh);
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

// entry: alive2 = alive on every tile, change flags (both parities) = the
// tile holds a revisit flag
template <bool kB>
__global__ void __launch_bounds__(kBlock)
    thin_tile_enter(vc_work W)
{
    for (uint32_t b = blockIdx.x; b < W.total; b += gridDim.x)
    {
        int vi;
        uint32_t lt;
        vc_item<kB>(W, b, vi, lt);
        const vc_vdesc d = vc_desc<kB>(W, vi);
        const int nwz = d.nwz;
        const int w = (int)(lt % (uint32_t)nwz);
        const uint32_t r2 = lt / (uint32_t)nwz;
        const int ty = (int)(r2 % (uint32_t)d.nty), tx = (int)(r2 / (uint32_t)d.nty);
        bool f = false;
        for (int q = threadIdx.x; q < kTT * kTT; q += kBlock)
        {
            const int x = tx * kTT + q / kTT, y = ty * kTT + q % kTT;
            if (x >= d.nx || y >= d.ny)
                continue;
            const int64_t o = ((int64_t)x * d.ny + y) * nwz + w;
            d.alive2[o] = d.alive[o];
            f = f || d.d0[o] || d.d1[o];
        }
        const bool any = __syncthreads_or(f);
        if (threadIdx.x == 0)
        {
            // a tile with new candidates: itself and every neighbour
            const int64_t nt = (int64_t)d.ntx * d.nty * nwz;
            d.tchg[lt] = any ? kAllNbrs : 0u;
            d.tchg[nt + lt] = any ? kAllNbrs : 0u;
        }
    }
}

// exclusive prefix sum over a tile-sweep workgroup (kTBlock threads) of
// v < 64: the wavefront's prefix from one ballot per bit plane of v (no
// cross-lane data movement), then the wave totals through wsum (kTBlock/64
// words; callers alternate two arrays so that no trailing barrier is needed
// before the next scan)
__device__ __forceinline__ uint32_t tile_scan(uint32_t v, uint32_t* wsum,
                                              uint32_t* total)
{
    const int wv = threadIdx.x >> 6;
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int bit = 0; bit < 6; ++bit)
    {
        const uint64_t m = __ballot((v >> bit) & 1u);
        pre += __builtin_amdgcn_mbcnt_hi(
                   (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))
               << bit;
        tot += (uint32_t)__popcll(m) << bit;
    }
    if ((threadIdx.x & 63) == 0)
        wsum[wv] = tot;
    __syncthreads();
    uint32_t base = 0, all = 0;
#pragma unroll
    for (int i = 0; i < kTBlock / 64; ++i)
    {
        const uint32_t t = wsum[i];
        base += i < wv ? t : 0u;
        all += t;
    }
    *total = all;
    return base + pre;
}

// One sweep over the visited tiles.  Each pass: the threads filter their
// rows' candidates (alive, eligible, z parity, range, a background face
// neighbour), queue them in LDS (row << 7 | bit), and every thread then takes
// kTQ / kTBlock queue entries: neighbourhood code from the LDS rows, table
// lookup, and on a hit an LDS atomic clears the bit.  Same-subfield voxels
// are not 26-adjacent, so a code never reads a bit that this pass clears.
template <bool kB>
__global__ void __launch_bounds__(kTBlock, kTWaves)
    thin_tile_sweep(vc_work W, const uint32_t* __restrict__ lut, int j,
                    int k_arg, int nvol, const uint32_t* __restrict__ ctl,
                    uint32_t* gany, uint32_t* vany,
                    unsigned long long* __restrict__ ctr)
{
    const PassCtl pc = pass_ctl(ctl, j, k_arg, 0, nvol, gany, vany);
    const uint32_t sweep = pc.sweep;
    const int k = pc.k;
    __shared__ unsigned long long SAL[kTRows];  // alive, window bits 0..63
    __shared__ unsigned long long SElo[kTRows]; // eligible (constant in a level)
    __shared__ uint32_t SH4[kTRows / 4];        // per row byte: alive bits 64..67
                                                // (low nibble), eligible (high)
    uint8_t* const SHB = reinterpret_cast<uint8_t*>(SH4);
    auto row = [&](int r) -> TRow {
        return TRow{SAL[r], (uint32_t)(SHB[r] & 0xFu), 0u};
    };
    __shared__ uint32_t tq[kTQ];
    __shared__ uint32_t red[2 * (kTBlock / 64)]; // scan totals, by pass parity
    __shared__ uint32_t snbr; // neighbour mask of this tile's removals
    // XCD-mapped: the tiles of one row column (consecutive items) and their
    // neighbours share rows, so they go to the same XCD's L2
    const uint32_t per = (W.total + 7) / 8;
    for (uint32_t L = blockIdx.x; L < 8 * per; L += gridDim.x)
    {
        const uint32_t b = vc_xcd_item(L, per);
        if (b >= W.total)
            continue;
        int vi;
        uint32_t lt;
        vc_item<kB>(W, b, vi, lt);
        const vc_vdesc d = vc_desc<kB>(W, vi);
        if (sweep > 0 && vany[4 * vi + ((sweep - 1) & 3)] == 0)
            continue; // converged at this level (block-uniform)
        const int nwz = d.nwz;
        const int64_t nt = (int64_t)d.ntx * d.nty * nwz;
        const int w = (int)(lt % (uint32_t)nwz);
        const uint32_t r2 = lt / (uint32_t)nwz;
        const int ty = (int)(r2 % (uint32_t)d.nty), tx = (int)(r2 / (uint32_t)d.nty);
        const uint32_t* cprev = d.tchg + (int64_t)((sweep + 1) & 1) * nt;
        uint32_t* ccur = d.tchg + (int64_t)(sweep & 1) * nt;
        // visit iff a tile of the 3x3x3 neighbourhood removed a voxel of this
        // tile's region in the last sweep: a tile whose region is unchanged
        // and that removed nothing would remove nothing again.  The first
        // sweep of a level visits every tile: it writes both planes' tiles
        // (reads alive, writes alive2), which stands in for the entry copy,
        // and its masks start the chain
        bool act = sweep == 0;
        if (sweep > 0 && threadIdx.x < 27)
        {
            const int ax = tx + (int)threadIdx.x / 9 - 1;
            const int ay = ty + ((int)threadIdx.x / 3) % 3 - 1;
            const int aw = w + (int)threadIdx.x % 3 - 1;
            if ((unsigned)ax < (unsigned)d.ntx && (unsigned)ay < (unsigned)d.nty &&
                (unsigned)aw < (unsigned)nwz)
                act = (cprev[((int64_t)ax * d.nty + ay) * nwz + aw] >>
                       (26 - threadIdx.x)) & 1u;
        }
        if (!__syncthreads_or(act))
        {
            if (threadIdx.x == 0)
                ccur[lt] = 0;
            continue;
        }
        if (threadIdx.x == 0)
            snbr = 0u;
        const uint64_t* Pin = (sweep & 1) ? d.alive2 : d.alive;
        uint64_t* Pout = (sweep & 1) ? d.alive : d.alive2;
        const int X0 = tx * kTT - kHX, Y0 = ty * kTT - kHY; // region origin
        for (int r = threadIdx.x; r < kTRows; r += kTBlock)
        {
            const int x = X0 + r / kRY, y = Y0 + r % kRY;
            uint64_t alo = 0, elo = 0;
            uint32_t ahi = 0, ehi = 0;
            if ((unsigned)x < (unsigned)d.nx && (unsigned)y < (unsigned)d.ny)
            {
                const int64_t o = ((int64_t)x * d.ny + y) * nwz + w;
                load_win(Pin, o, w, nwz, alo, ahi);
                load_win(d.elig, o, w, nwz, elo, ehi);
            }
            SAL[r] = alo;
            SElo[r] = elo;
            SHB[r] = (uint8_t)(ahi | (ehi << 4));
        }
        __syncthreads();
        uint32_t nrem = 0, ntest = 0, nbr = 0;
        uint8_t* lrow0 = d.label + (int64_t)64 * w;
        for (int p = 0; p < 8; ++p)
        {
            const int sx = p & 1, sy = (p >> 1) & 1, sz = (p >> 2) & 1;
            const int Fx = 7 - p, Fy = fy_of(p), Fz = p < 4 ? 1 : 0;
            // rows [kHX - Fx, kHX + kTT + Fx) x [kHY - Fy, kHY + kTT + Fy) of
            // parity (sx, sy); the region origin is even, so a row's parity
            // is its region index's
            const int xlo = kHX - Fx, xhi = kHX + kTT + Fx;
            const int ylo = kHY - Fy, yhi = kHY + kTT + Fy;
            const int rx0 = xlo + ((xlo & 1) != sx), ry0 = ylo + ((ylo & 1) != sy);
            const int nrx = (xhi - rx0 + 1) >> 1, nry = (yhi - ry0 + 1) >> 1;
            // window bit parity = z parity (64w - 2 is even); bits
            // [kHZ - Fz, kHZ + 64 + Fz): low word from bit 1 or 2 up, high
            // bits 64..65 or 64..66
            const uint64_t zs64 = sz ? 0xAAAAAAAAAAAAAAAAull : 0x5555555555555555ull;
            const uint64_t zlo = zs64 & (~0ull << (kHZ - Fz));
            const uint32_t zhi = (uint32_t)zs64 & ((1u << (kHZ + Fz)) - 1u);
            uint64_t clo[kTRowsPerThread];
            uint32_t chi[kTRowsPerThread];
            int rr[kTRowsPerThread];
            uint32_t cnt = 0;
#pragma unroll
            for (int t = 0; t < kTRowsPerThread; ++t)
            {
                clo[t] = 0;
                chi[t] = 0;
                rr[t] = 0;
                const int q = threadIdx.x + t * kTBlock;
                if (q >= nrx * nry)
                    continue;
                const int rx = rx0 + 2 * (q / nry), ry = ry0 + 2 * (q % nry);
                const int r = rx * kRY + ry;
                const uint32_t hb = SHB[r];
                const TRow C{SAL[r], hb & 0xFu, 0u};
                uint64_t cl = C.lo & SElo[r] & zlo;
                uint32_t ch = C.hi & (hb >> 4) & zhi;
                if (cl | ch)
                {
                    // border filter: all six face neighbours alive -> never
                    // simple
                    const TRow a = row(r - kRY), bb = row(r + kRY), c = row(r - 1),
                               e = row(r + 1);
                    const uint64_t zdn = (C.lo << 1);
                    const uint64_t zup = (C.lo >> 1) | ((uint64_t)C.hi << 63);
                    cl &= ~(a.lo & bb.lo & c.lo & e.lo & zdn & zup);
                    const uint32_t hdn = (C.hi << 1) | (uint32_t)(C.lo >> 63);
                    const uint32_t hup = C.hi >> 1;
                    ch &= ~(a.hi & bb.hi & c.hi & e.hi & hdn & hup);
                }
                clo[t] = cl;
                chi[t] = ch;
                rr[t] = r;
                cnt += (uint32_t)(__popcll(cl) + __popc(ch));
            }
            uint32_t total;
            uint32_t idx = tile_scan(cnt, red + (p & 1) * (kTBlock / 64), &total);
            ntest += total;
            for (uint32_t c0 = 0; c0 < total; c0 += kTQ)
            {
#pragma unroll
                for (int t = 0; t < kTRowsPerThread; ++t)
                {
                    const uint32_t base = (uint32_t)rr[t] << 7;
                    while (clo[t] && idx < c0 + kTQ)
                    {
                        const int i = __ffsll((unsigned long long)clo[t]) - 1;
                        clo[t] &= clo[t] - 1;
                        tq[idx - c0] = base | (uint32_t)i;
                        ++idx;
                    }
                    while (chi[t] && idx < c0 + kTQ)
                    {
                        const int i = __ffs(chi[t]) - 1;
                        chi[t] &= chi[t] - 1;
                        tq[idx - c0] = base | (uint32_t)(64 + i);
                        ++idx;
                    }
                }
                __syncthreads();
                const uint32_t n = min((uint32_t)kTQ, total - c0);
                uint32_t ent[kTQ / kTBlock], code[kTQ / kTBlock],
                    lw[kTQ / kTBlock];
#pragma unroll
                for (int t = 0; t < kTQ / kTBlock; ++t)
                {
                    const uint32_t e = threadIdx.x + t * kTBlock;
                    ent[t] = e < n ? tq[e] : 0xFFFFFFFFu;
                    code[t] = 0;
                    if (e < n)
                    {
                        const int r = (int)(ent[t] >> 7);
                        TRow R[3][3];
#pragma unroll
                        for (int dx = 0; dx < 3; ++dx)
#pragma unroll
                            for (int dy = 0; dy < 3; ++dy)
                                R[dx][dy] = row(r + (dx - 1) * kRY + (dy - 1));
                        code[t] = nb_code_rows(R, (int)(ent[t] & 127u));
                    }
                }
#pragma unroll
                for (int t = 0; t < kTQ / kTBlock; ++t)
                    lw[t] = ent[t] != 0xFFFFFFFFu ? lut[code[t] >> 5] : 0u;
#pragma unroll
                for (int t = 0; t < kTQ / kTBlock; ++t)
                {
                    if (ent[t] == 0xFFFFFFFFu || !((lw[t] >> (code[t] & 31)) & 1u))
                        continue;
                    const int r = (int)(ent[t] >> 7), i = (int)(ent[t] & 127u);
                    if (i < 64)
                        atomicAnd(&SAL[r], ~(1ull << i));
                    else
                        atomicAnd(&SH4[r >> 2], ~(1u << (8 * (r & 3) + (i - 64))));
                    const int rx = r / kRY, ry = r % kRY;
                    if (rx >= kHX && rx < kHX + kTT && ry >= kHY &&
                        ry < kHY + kTT && i >= kHZ && i < kHZ + 64)
                    {
                        ++nrem;
                        nbr |= halo_nbrs(rx - kHX, ry - kHY, i - kHZ);
                        lrow0[((int64_t)(X0 + rx) * d.ny + (Y0 + ry)) * d.nz +
                              (i - kHZ)] = (uint8_t)(2 + k);
                    }
                }
                __syncthreads();
            }
            // no barrier here: a pass with candidates ended with its last
            // round's barrier (after the atomics the next pass reads); one
            // without changed nothing, and the next scan uses the other red
        }
        for (int q = threadIdx.x; q < kTT * kTT; q += kTBlock)
        {
            const int rx = kHX + q / kTT, ry = kHY + q % kTT;
            const int x = X0 + rx, y = Y0 + ry;
            if (x < d.nx && y < d.ny)
            {
                const TRow t = row(rx * kRY + ry);
                Pout[((int64_t)x * d.ny + y) * nwz + w] =
                    (t.lo >> kHZ) | ((uint64_t)t.hi << (64 - kHZ));
            }
        }
#pragma unroll
        for (int dd = 32; dd >= 1; dd >>= 1)
            nbr |= __shfl_xor(nbr, dd);
        if ((threadIdx.x & 63) == 0 && nbr)
            atomicOr(&snbr, nbr);
        const uint32_t tot = block_sum(nrem, red); // orders the atomics
        if (threadIdx.x == 0)
        {
            ccur[lt] = snbr;
            unsigned long long* c =
                ctr + (size_t)(b % VC_CTR_SLOTS) * VC_CTR_WORDS;
            if (tot)
            {
                gany[sweep & (VC_RING - 1)] = 1u;
                vany[4 * vi + (sweep & 3)] = 1u;
                atomicAdd(&c[8 + k], (unsigned long long)tot);
            }
            if (ntest)
                atomicAdd(&c[0], (unsigned long long)ntest);
            // the tile's own voxels, once per subfield pass
            const int64_t tv = (int64_t)min(kTT, d.nx - tx * kTT) *
                               min(kTT, d.ny - ty * kTT) * min(64, d.nz - 64 * w);
            atomicAdd(&c[1], (unsigned long long)(8 * tv));
        }
        __syncthreads(); // the LDS rows, tq, red are rewritten by the next item
    }
}

// the next segment's first sweep
__global__ void seg_advance(uint32_t* ctl, int K)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
        ctl[0] += (uint32_t)K;
}

int pass_count(const vc_sched* s)
{
    int c = 0;
    for (int sub = 0; sub < 8; ++sub)
        c += s->total[VK_TP + (sub & 3)] != 0;
    return c;
}

// K sweeps (8 subfield passes each).  Captured (graph): sweep j of the
// segment reads the segment's first sweep and the level from ctl, and the
// segment ends by advancing ctl[0]; launched directly: the absolute sweep
// (first + j) and the level are arguments.
template <bool kB, bool kTail>
void launch_pass(hipStream_t S, unsigned g, const vc_work& W,
                 const uint32_t* lut, int sub, int j, int k, int n,
                 const uint32_t* ctl, vc_sched* s)
{
    if (kTail)
        thin_pass_sparse<kB><<<g, 64, 0, S>>>(W, lut, sub, j, k, n, ctl,
                                              s->d_gany, s->d_vany, s->d_ctr);
    else
        thin_pass<kB><<<g, kBlock, 0, S>>>(W, lut, sub, j, k, n, ctl,
                                           s->d_gany, s->d_vany, s->d_ctr);
}

// sweeps with index >= tail_sweep within a level use the sparse variant
bool tail_sweep_of(const vc_sched* s, int64_t sweep)
{
    return s->sc.tail_sweep >= 0 && sweep >= s->sc.tail_sweep;
}

// how a segment's sweeps run
enum SegMode
{
    SEG_DENSE = 0,
    SEG_SPARSE = 1,
    SEG_TILE = 2,       // tile sweeps
    SEG_TILE_ENTRY = 3  // tile sweeps, entry kernel first
};

template <bool kB>
void launch_tile_sweep(hipStream_t S, unsigned g, const vc_work& W,
                       const uint32_t* lut, int j, int k, int n,
                       const uint32_t* ctl, vc_sched* s)
{
    thin_tile_sweep<kB><<<g, kTBlock, 0, S>>>(W, lut, j, k, n, ctl, s->d_gany,
                                             s->d_vany, s->d_ctr);
}

// sweep timing: an event from the context's pool (nullptr when none can be
// created)
hipEvent_t ev_take(vc_ctx* ctx)
{
    if (ctx->ev_free.empty())
    {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess)
            return nullptr;
        return e;
    }
    hipEvent_t e = ctx->ev_free.back();
    ctx->ev_free.pop_back();
    return e;
}

// records the first event of a timed region; e[0] = e[1] = nullptr when
// timing is off
int ev_begin(vc_ctx* ctx, hipEvent_t* e)
{
    e[0] = e[1] = nullptr;
    if (!ctx->time_sweeps)
        return VC_OK;
    e[0] = ev_take(ctx);
    e[1] = ev_take(ctx);
    if (!e[0] || !e[1])
    {
        for (int i = 0; i < 2; ++i)
            if (e[i])
                ctx->ev_free.push_back(e[i]);
        return vc_fail(ctx, VC_ERR_HIP, "sweep timing: hipEventCreate failed");
    }
    VC_HIP(ctx, hipEventRecord(e[0], ctx->stream));
    return VC_OK;
}

int ev_end(vc_ctx* ctx, hipEvent_t* e, int launches)
{
    if (!e[0])
        return VC_OK;
    VC_HIP(ctx, hipEventRecord(e[1], ctx->stream));
    ctx->ev_open.emplace_back(e[0], e[1]);
    ctx->ev_open_n.push_back(launches);
    return VC_OK;
}

// after a host wait on the context stream: add the closed regions
int ev_collect(vc_ctx* ctx)
{
    for (size_t i = 0; i < ctx->ev_open.size(); ++i)
    {
        float ms = 0.0f;
        const hipError_t e = hipEventElapsedTime(&ms, ctx->ev_open[i].first,
                                                 ctx->ev_open[i].second);
        if (e != hipSuccess)
            return vc_fail(ctx, VC_ERR_HIP, "sweep timing: %s",
                           hipGetErrorString(e));
        ctx->sweep_ms += ms;
        ctx->sweep_launches += ctx->ev_open_n[i];
        ctx->ev_free.push_back(ctx->ev_open[i].first);
        ctx->ev_free.push_back(ctx->ev_open[i].second);
    }
    ctx->ev_open.clear();
    ctx->ev_open_n.clear();
    return VC_OK;
}

int launch_segment(vc_ctx* ctx, vc_sched* s, int K, bool graph, int64_t first,
                   int k, int mode)
{
    hipStream_t S = ctx->stream;
    const int pc = s->sc.thin_grid_per_cu;
    const uint32_t* ctl = graph ? s->d_ctl : nullptr;
    if (mode >= SEG_TILE)
    {
        const unsigned g = vc_sched_grid(ctx, s, VK_TT, pc);
        const vc_work W = vc_sched_work(s, VK_TT);
        // tile sweeps from the level's first sweep need no entry kernel: that
        // sweep visits every tile (see thin_tile_sweep)
        if (g && mode == SEG_TILE_ENTRY && s->sc.tile_sweep != 0)
        {
            if (s->n > 1)
                thin_tile_enter<true><<<g, kBlock, 0, S>>>(W);
            else
                thin_tile_enter<false><<<g, kBlock, 0, S>>>(W);
            VC_CHECK_LAUNCH(ctx);
        }
        for (int j = 0; j < K && g; ++j)
        {
            const int jj = graph ? j : (int)(first + j);
            hipEvent_t ev[2] = {nullptr, nullptr};
            if (!graph)
            {
                const int rc = ev_begin(ctx, ev);
                if (rc != VC_OK)
                    return rc;
            }
            if (s->n > 1)
                launch_tile_sweep<true>(S, g, W, ctx->d_lut, jj, k, s->n, ctl, s);
            else
                launch_tile_sweep<false>(S, g, W, ctx->d_lut, jj, k, s->n, ctl,
                                         s);
            VC_CHECK_LAUNCH(ctx);
            const int rc = ev_end(ctx, ev, 1);
            if (rc != VC_OK)
                return rc;
        }
    }
    else
        for (int j = 0; j < K; ++j)
        {
            const bool tail =
                graph ? mode == SEG_SPARSE : tail_sweep_of(s, first + j);
            for (int sub = 0; sub < 8; ++sub)
            {
                const int kind = VK_TP + (sub & 3);
                const unsigned g = vc_sched_grid(ctx, s, kind, pc);
                if (!g)
                    continue;
                const vc_work W = vc_sched_work(s, kind);
                const int jj = graph ? j : (int)(first + j);
                if (s->n > 1)
                {
                    if (tail)
                        launch_pass<true, true>(S, g, W, ctx->d_lut, sub, jj, k,
                                                s->n, ctl, s);
                    else
                        launch_pass<true, false>(S, g, W, ctx->d_lut, sub, jj,
                                                 k, s->n, ctl, s);
                }
                else if (tail)
                    launch_pass<false, true>(S, g, W, ctx->d_lut, sub, jj, k,
                                             s->n, ctl, s);
                else
                    launch_pass<false, false>(S, g, W, ctx->d_lut, sub, jj, k,
                                              s->n, ctl, s);
                VC_CHECK_LAUNCH(ctx);
            }
        }
    if (graph)
    {
        seg_advance<<<1, 64, 0, S>>>(s->d_ctl, K);
        VC_CHECK_LAUNCH(ctx);
    }
    return VC_OK;
}

// a segment as a captured graph (captured once per length and mode, then
// replayed); dense/sparse graph segments use one variant, chosen by their
// first sweep
int run_segment(vc_ctx* ctx, vc_sched* s, int K, int64_t first, int k,
                int mode)
{
    if (!s->sc.use_graph)
        return launch_segment(ctx, s, K, false, first, k, mode);
    hipStream_t S = ctx->stream;
    const int key = K * 4 + mode;
    auto it = s->graphs.find(key);
    if (it == s->graphs.end())
    {
        VC_HIP(ctx, hipStreamBeginCapture(S, hipStreamCaptureModeThreadLocal));
        const int rc = launch_segment(ctx, s, K, true, 0, 0, mode);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(S, &g);
        if (rc != VC_OK)
        {
            if (g)
                (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess)
            return vc_fail(ctx, VC_ERR_HIP, "segment capture: %s",
                           hipGetErrorString(e));
        hipGraphExec_t ex = nullptr;
        const hipError_t e2 = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e2 != hipSuccess)
            return vc_fail(ctx, VC_ERR_HIP, "segment instantiate: %s",
                           hipGetErrorString(e2));
        it = s->graphs.emplace(key, ex).first;
    }
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (mode >= SEG_TILE)
    {
        const int rc = ev_begin(ctx, ev);
        if (rc != VC_OK)
            return rc;
    }
    VC_HIP(ctx, hipGraphLaunch(it->second, S));
    return ev_end(ctx, ev, K);
}

int upload_levels(vc_ctx* ctx, vc_sched* s, const float* levels, int nlev)
{
    const size_t nl = (size_t)s->n * nlev;
    if (s->lv_n == nlev && s->lv_cache.size() == nl &&
        std::memcmp(s->lv_cache.data(), levels, 4 * nl) == 0)
        return VC_OK;
    // the staging buffer may still feed an earlier copy
    VC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const size_t cells = (size_t)s->n * VC_MAX_LEVELS;
    uint64_t* hT = static_cast<uint64_t*>(s->h_lv);
    float* hL = reinterpret_cast<float*>(hT + cells);
    std::memset(s->h_lv, 0, 12 * cells);
    for (int v = 0; v < s->n; ++v)
        for (int q = 0; q < nlev; ++q)
        {
            const float l = levels[(size_t)v * nlev + q];
            hT[(size_t)v * VC_MAX_LEVELS + q] = (uint64_t)vc_level_threshold(l);
            hL[(size_t)v * VC_MAX_LEVELS + q] = l;
        }
    VC_HIP(ctx, hipMemcpyAsync(s->d_T, hT, 8 * cells, hipMemcpyHostToDevice,
                               ctx->stream));
    VC_HIP(ctx, hipMemcpyAsync(s->d_L, hL, 4 * cells, hipMemcpyHostToDevice,
                               ctx->stream));
    s->lv_cache.assign(levels, levels + nl);
    s->lv_n = nlev;
    return VC_OK;
}

} // namespace

int vc_thin_timing(vc_ctx* ctx, int enable)
{
    if (!ctx || (enable != 0 && enable != 1))
        return VC_ERR_INVALID;
    VC_HIP(ctx, hipSetDevice(ctx->device));
    VC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (auto& pr : ctx->ev_open)
    {
        ctx->ev_free.push_back(pr.first);
        ctx->ev_free.push_back(pr.second);
    }
    ctx->ev_open.clear();
    ctx->ev_open_n.clear();
    ctx->time_sweeps = enable != 0;
    ctx->sweep_launches = 0;
    ctx->sweep_ms = 0.0;
    return VC_OK;
}

int vc_thin_timing_read(vc_ctx* ctx, int64_t* launches, double* total_ms)
{
    if (!ctx || !launches || !total_ms)
        return VC_ERR_INVALID;
    *launches = ctx->sweep_launches;
    *total_ms = ctx->sweep_ms;
    return VC_OK;
}

int vc_lut_build(vc_ctx* ctx)
{
    int rc = vc_alloc(ctx, (void**)&ctx->d_lut, sizeof(uint32_t) << 21);
    if (rc != VC_OK)
        return rc;
    lut_build<<<(1u << 21) / kBlock, kBlock, 0, ctx->stream>>>(ctx->d_lut);
    VC_CHECK_LAUNCH(ctx);
    return VC_OK;
}

int vc_thin_sched(vc_ctx* ctx, vc_sched* s, const float* levels, int nlev,
                  vc_thin_stats* st)
{
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    int rc = vc_sched_sync_measure(ctx, s);
    if (rc == VC_OK)
        rc = upload_levels(ctx, s, levels, nlev);
    if (rc != VC_OK)
        return rc;
    const size_t ctr_bytes = 8 * (size_t)VC_CTR_SLOTS * VC_CTR_WORDS;
    VC_HIP(ctx, hipMemsetAsync(s->d_ctr, 0, ctr_bytes, S));
    const vc_schedule& sc = s->sc;
    const int per_seg = pass_count(s) + 1;
    int64_t launches = 0, passes_total = 0;
    // every sweep before convergence removes a voxel; sweep indices are int
    // kernel arguments
    const int64_t guard =
        s->voxels + 2 < (int64_t)0x7FFFFFFF - 64 ? s->voxels + 2
                                                 : (int64_t)0x7FFFFFFF - 64;
    for (int k = 0; k < nlev; ++k)
    {
        const int li_kind = k == 0 ? VK_LI : VK_LN;
        const unsigned g = vc_sched_grid(ctx, s, li_kind, sc.thin_grid_per_cu);
        if (g)
        {
            const vc_work W = vc_sched_work(s, li_kind);
            if (k > 0)
            {
                if (s->n > 1)
                    thin_level_next<true><<<g, kBlock, 0, S>>>(
                        W, k, s->n, s->d_ctl, s->d_gany, s->d_vany);
                else
                    thin_level_next<false><<<g, kBlock, 0, S>>>(
                        W, k, s->n, s->d_ctl, s->d_gany, s->d_vany);
            }
            else if (s->n > 1)
                thin_level_init<true><<<g, kBlock, 0, S>>>(
                    W, s->d_T, s->d_L, nlev, k, s->n, s->d_ctl, s->d_gany,
                    s->d_vany);
            else
                thin_level_init<false><<<g, kBlock, 0, S>>>(
                    W, s->d_T, s->d_L, nlev, k, s->n, s->d_ctl, s->d_gany,
                    s->d_vany);
            VC_CHECK_LAUNCH(ctx);
            launches++;
        }
        int64_t base = 0;
        int K = sc.sweeps_per_sync > 0 ? sc.sweeps_per_sync : VC_SEG_MIN;
        bool in_tile = false;
        for (;;)
        {
            if (base > guard)
                return vc_fail(ctx, VC_ERR_GUARD,
                               "thinning did not converge in %lld sweeps",
                               (long long)base);
            // tile sweeps from the first segment starting at >= tile_sweep
            const bool tile = sc.tile_sweep >= 0 && base >= sc.tile_sweep &&
                              s->total[VK_TT] != 0;
            const int mode = tile ? (in_tile ? SEG_TILE : SEG_TILE_ENTRY)
                                  : (tail_sweep_of(s, base) ? SEG_SPARSE
                                                            : SEG_DENSE);
            rc = run_segment(ctx, s, K, base, k, mode);
            if (rc != VC_OK)
                return rc;
            if (tile)
                launches += (int64_t)K + (in_tile || sc.tile_sweep == 0 ? 0 : 1) +
                            (sc.use_graph ? 1 : 0);
            else
                launches += (int64_t)K * (per_seg - 1) + (sc.use_graph ? 1 : 0);
            in_tile = tile;
            VC_HIP(ctx, hipMemcpyAsync(s->h_ring, s->d_gany, 4 * VC_RING,
                                       hipMemcpyDeviceToHost, S));
            VC_HIP(ctx, hipStreamSynchronize(S));
            rc = ev_collect(ctx);
            if (rc != VC_OK)
                return rc;
            int64_t conv = -1;
            for (int jj = 0; jj < K; ++jj)
                if (s->h_ring[(base + jj) & (VC_RING - 1)] == 0)
                {
                    conv = base + jj;
                    break;
                }
            base += K;
            if (conv >= 0)
            {
                passes_total += 8 * (conv + 1);
                break;
            }
            // adaptive segments: by default they grow by VC_SEG_MIN sweeps
            // (4, 8, 12, 16) — a level converges inside its last segment,
            // whose remaining sweeps launch and find nothing, so arithmetic
            // growth overshoots less than doubling (512^3 headline: 110
            // launches instead of 122, -1.6 %); sweeps_per_sync = -1 doubles
            // (4, 8, 16: fewer host checks for levels of ~25 sweeps)
            if (sc.sweeps_per_sync == 0)
                K = K + VC_SEG_MIN < VC_SEG_MAX ? K + VC_SEG_MIN : VC_SEG_MAX;
            else if (sc.sweeps_per_sync < 0)
                K = K * 2 < VC_SEG_MAX ? K * 2 : VC_SEG_MAX;
        }
    }
    if (st)
    {
        VC_HIP(ctx, hipMemcpyAsync(s->h_ctr, s->d_ctr, ctr_bytes,
                                   hipMemcpyDeviceToHost, S));
        VC_HIP(ctx, hipStreamSynchronize(S));
        std::memset(st, 0, sizeof(*st));
        for (int sl = 0; sl < VC_CTR_SLOTS; ++sl)
        {
            const unsigned long long* c = s->h_ctr + (size_t)sl * VC_CTR_WORDS;
            st->tests += (int64_t)c[0];
            st->voxel_passes += (int64_t)c[1];
            for (int q = 0; q < nlev; ++q)
                st->removed[q] += (int64_t)c[8 + q];
        }
        st->passes = passes_total;
        st->launches = launches;
    }
    for (vc_vol* v : s->vols)
        v->thin_done = true;
    return VC_OK;
}
