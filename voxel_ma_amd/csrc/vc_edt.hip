// vc_edt.hip -- exact squared Euclidean distance transform + feature transform
// and the per-voxel medial measure, gfx950.
//
// Three separable passes, each one read + one write of HBM:
//   pass z (contiguous rows):  uint8 occupancy -> int16 nearest background z'
//   pass y (columns, stride nz): int16 z'      -> uint32 packed (y', z')
//   pass x (columns, stride ny*nz): (y', z')   -> uint32 packed FT + int32 d2
// Algorithmic bytes per voxel: (1+2) + (2+4) + (4+4+4) = 21.
//
// The row pass finds the nearest background voxel from a per-row background
// bitmask built with wavefront ballots in LDS.  The column passes stage a
// tile of TZ consecutive z columns in LDS (coalesced TZ-wide row segments)
// and resolve each voxel by an outward search of the lower envelope:
// candidates at distance d cost >= d^2, so the search stops once d^2 exceeds
// the best value; ties go to the lower index, which makes the final feature
// the lowest linear index among all nearest background voxels.
//
// Every kernel walks a work table (one volume or a batch of volumes, one
// launch per pass over all of them); with a persistent grid each workgroup
// loops over its share of the items.
#include <algorithm>

#include "vc_internal.h"

namespace
{

constexpr int kBlock = 256;
constexpr int kRowsZ = 4; // pass z: one wavefront per row, 4 rows per item
constexpr uint32_t kOob = 0xFFFFFFFEu; // neighbour outside the grid
// every candidate cost fits in int32: the volume constructor rejects shapes
// whose squared diagonal exceeds INT32_MAX
constexpr int kInf = 0x7FFFFFFF;

// Thick solids.  A column pass's outward search for a voxel ends within the
// distance of the voxel's own staged candidate (its cost bounds the best),
// which the previous pass knows: pass z flags the y-pass tiles holding a
// voxel more than kDeepY from the background along z (or with none in its
// row), pass y the x-pass tiles holding a voxel whose (y, z) cost exceeds
// kDeepX^2 (or has none).  A flagged tile is computed by divide and conquer
// (column_argmins), whose work does not grow with the distance, and clears
// its flag; the others run the outward search.  Either way the result is the
// leftmost minimiser, so a stale flag costs time, never correctness.  On the
// 512^3 blurred-noise volume no tile is flagged (its longest z run of inside
// voxels has a half-length below 128; at a y threshold of 64, 13 % of the
// y-pass tiles were, and the EDT took 1.44 instead of 1.26 ms); in a ball of
// radius 250 most are (EDT 10.2 -> 4.1 ms).
constexpr int kDeepY = 128, kDeepX = 48;

// rows along z: wavefront = row; the row's background bitmask sits in the
// wavefront's LDS slice of `stride` words (rows whose length is a multiple
// of 8: 8 voxels per lane and access; others: one ballot per 64 voxels)
template <bool kB>
__global__ void __launch_bounds__(kBlock)
    edt_pass_z(vc_work W, int stride)
{
    extern __shared__ uint64_t lds_z[];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    uint64_t* bgmask = lds_z + (size_t)wv * stride;
    for (uint32_t b = blockIdx.x; b < W.total; b += gridDim.x)
    {
        int vi;
        uint32_t lt;
        vc_item<kB>(W, b, vi, lt);
        const vc_vdesc d = vc_desc<kB>(W, vi);
        const int nz = d.nz;
        const int64_t row = (int64_t)lt * kRowsZ + wv;
        const bool active = row < (int64_t)d.nx * d.ny;
        const uint8_t* r = d.occ + row * nz;
        const int nw = (nz + 63) >> 6;
        if ((nz & 7) == 0)
        {
            // 8 voxels per lane: one 8-byte load, a background byte by SWAR
            // zero-byte detection, one nearest-background search per side of
            // the lane's block, one 16-byte store
            uint8_t* bgb = reinterpret_cast<uint8_t*>(bgmask);
            if (active)
            {
                // the bytes of the last word past the row end hold no
                // background
                for (int q = (nz >> 3) + lane; q < nw * 8; q += 64)
                    bgb[q] = 0;
                for (int z0 = lane * 8; z0 < nz; z0 += 512)
                {
                    const uint64_t v =
                        *reinterpret_cast<const uint64_t*>(r + z0);
                    constexpr uint64_t k7f = 0x7F7F7F7F7F7F7F7Full;
                    const uint64_t zb = ~(((v & k7f) + k7f) | v | k7f);
                    bgb[z0 >> 3] = (uint8_t)(((zb >> 7) * 0x0102040810204080ull) >> 56);
                }
            }
            __syncthreads();
            if (active)
            {
                int16_t* out = reinterpret_cast<int16_t*>(d.edt) + row * nz;
                for (int z0 = lane * 8; z0 < nz; z0 += 512)
                {
                    const int w = z0 >> 6, k = (z0 >> 3) & 7;
                    const uint64_t m = bgmask[w];
                    const uint32_t b = (uint32_t)(m >> (8 * k)) & 0xFFu;
                    int Lb = -1, Rb = -1;
                    const uint64_t lm = k ? m & ((1ull << (8 * k)) - 1ull) : 0ull;
                    if (lm)
                        Lb = (w << 6) + 63 - __clzll(lm);
                    else
                        for (int ww = w - 1; ww >= 0; --ww)
                        {
                            const uint64_t q = bgmask[ww];
                            if (q)
                            {
                                Lb = (ww << 6) + 63 - __clzll(q);
                                break;
                            }
                        }
                    const uint64_t rm = k < 7 ? m & ~((1ull << (8 * k + 8)) - 1ull) : 0ull;
                    if (rm)
                        Rb = (w << 6) + __ffsll((unsigned long long)rm) - 1;
                    else
                        for (int ww = w + 1; ww < nw; ++ww)
                        {
                            const uint64_t q = bgmask[ww];
                            if (q)
                            {
                                Rb = (ww << 6) + __ffsll((unsigned long long)q) - 1;
                                break;
                            }
                        }
                    uint32_t packed[4];
#pragma unroll
                    for (int j = 0; j < 8; ++j)
                    {
                        const int z = z0 + j;
                        int res;
                        if ((b >> j) & 1u)
                            res = z;
                        else
                        {
                            const uint32_t below = b & ((1u << j) - 1u);
                            const uint32_t above = b & ~((2u << j) - 1u);
                            const int L = below ? z0 + 31 - __clz(below) : Lb;
                            const int R = above ? z0 + __ffs(above) - 1 : Rb;
                            res = L < 0 ? R : (R < 0 ? L : (z - L <= R - z ? L : R));
                            if (res < 0 || abs(z - res) > kDeepY)
                                d.fy[(row / d.ny) * ((nz + d.tzy - 1) / d.tzy) + z / d.tzy] = 1;
                        }
                        if (j & 1)
                            packed[j >> 1] |= (uint32_t)(uint16_t)res << 16;
                        else
                            packed[j >> 1] = (uint32_t)(uint16_t)res;
                    }
                    *reinterpret_cast<uint4*>(out + z0) =
                        make_uint4(packed[0], packed[1], packed[2], packed[3]);
                }
            }
            __syncthreads(); // the slice is rewritten by the next item
            continue;
        }
        if (active)
            for (int w = 0; w < nw; ++w)
            {
                const int z = (w << 6) + lane;
                const bool bg = (z < nz) && r[z] == 0;
                const uint64_t m = __ballot(bg);
                if (lane == 0)
                    bgmask[w] = m;
            }
        __syncthreads();
        if (active)
        {
            int16_t* out = reinterpret_cast<int16_t*>(d.edt) + row * nz;
            for (int z = lane; z < nz; z += 64)
            {
                const int w = z >> 6, bb = z & 63;
                const uint64_t m = bgmask[w];
                int res;
                if ((m >> bb) & 1ull)
                    res = z;
                else
                {
                    int L = -1, R = -1;
                    const uint64_t lm = m & ((1ull << bb) - 1ull); // below bb
                    if (lm)
                        L = (w << 6) + 63 - __clzll(lm);
                    else
                        for (int ww = w - 1; ww >= 0; --ww)
                        {
                            const uint64_t q = bgmask[ww];
                            if (q)
                            {
                                L = (ww << 6) + 63 - __clzll(q);
                                break;
                            }
                        }
                    const uint64_t rm = m & ~((2ull << bb) - 1ull); // above bb
                    if (rm)
                        R = (w << 6) + __ffsll((unsigned long long)rm) - 1;
                    else
                        for (int ww = w + 1; ww < nw; ++ww)
                        {
                            const uint64_t q = bgmask[ww];
                            if (q)
                            {
                                R = (ww << 6) + __ffsll((unsigned long long)q) - 1;
                                break;
                            }
                        }
                    if (L < 0)
                        res = R;
                    else if (R < 0)
                        res = L;
                    else
                        res = (z - L <= R - z) ? L : R;
                    if (res < 0 || abs(z - res) > kDeepY)
                        d.fy[(row / d.ny) * ((nz + d.tzy - 1) / d.tzy) + z / d.tzy] = 1;
                }
                out[z] = (int16_t)res;
            }
        }
        __syncthreads(); // the slice is rewritten by the next item
    }
}

// columns along y: item = TZ consecutive z at fixed x, all y.  z' comes from
// the first half of the edt buffer, (y', z') goes to the ft buffer.
//
// kM (every volume has ny, nz <= 16384): the staged element is the signed
// offset z - z' (int16), so a candidate's cost is one 24-bit multiply-add,
// (z - z')^2 + dy^2; a row without background is stored as -32768, whose
// cost 2^30 + dy^2 is at least kNoBg, above every real cost (< 2^29).
constexpr int kNoBg = 1 << 30;


// Leftmost minimum over an aligned group of G lanes (G a power of two, at
// most 64): the smallest cost, then the smallest index; index -1 = none.
__device__ __forceinline__ void group_min(int& c, int& q, int G)
{
    for (int d = 1; d < G; d <<= 1)
    {
        const int oc = __shfl_xor(c, d), oq = __shfl_xor(q, d);
        if (oq >= 0 && (q < 0 || oc < c || (oc == c && oq < q)))
        {
            c = oc;
            q = oq;
        }
    }
}

// Leftmost argmin A(y) of cost(y, q) = f(q) + (y - q)^2 over q in [0, n),
// for every y in [0, n) of the tile's columns c < ncol, by divide and
// conquer: cost is a Monge array ((y' - y)(q' - q) >= 0 orders the cross
// terms), so A is non-decreasing in y and A(mid) of a segment lies between
// the A of the segment's two bounding rows.  Rows are the nodes of the
// perfect binary tree over [0, 2^L - 1): level l holds 2^l nodes, node j has
// mid ((2j+1) << (L-1-l)) - 1 and bounding rows (j << (L-l)) - 1 and
// ((j+1) << (L-l)) - 1, mids of earlier levels (or outside [0, n)).  Each
// level spreads its (column, node) tasks over the workgroup, G lanes per
// task while tasks are fewer than lanes; about n log2 n cost evaluations per
// column.  A column without a finite cost gets -1 everywhere.
//   cost(c, y, q): the cost, >= lim when q has no background;
//   put(c, y, q, cost): the result of row y (q = -1: none), to global memory;
//   get(c, y): A(y) of a row put at an earlier level (-1: none), read back
//   after the level's barrier (global writes of a workgroup are visible to
//   the workgroup after __syncthreads).
// Every lane of the workgroup must call it (barriers, shuffles).
template <class Cost, class Put, class Get>
__device__ void column_argmins(int n, int ncol, int TZ, int lim, Cost cost,
                               Put put, Get get)
{
    int L = 1;
    while ((1 << L) - 1 < n)
        ++L;
    const int nthr = (int)blockDim.x;
    for (int l = 0; l < L; ++l)
    {
        const int ntask = TZ << l;
        int gl = 0;
        while (gl < 6 && (ntask << (gl + 1)) <= nthr)
            ++gl;
        const int G = 1 << gl, ngroups = nthr >> gl;
        const int sub = (int)threadIdx.x & (G - 1);
        const int trips = (ntask + ngroups - 1) / ngroups;
        for (int it = 0; it < trips; ++it)
        {
            const int task = it * ngroups + ((int)threadIdx.x >> gl);
            const int c = task % TZ, j = task / TZ;
            const int mid = ((2 * j + 1) << (L - 1 - l)) - 1;
            const bool active = task < ntask && c < ncol && mid < n;
            int bc = lim, bq = -1;
            if (active)
            {
                const int yl = (j << (L - l)) - 1, yr = ((j + 1) << (L - l)) - 1;
                int qlo = 0, qhi = n - 1;
                bool none = false;
                if (yl >= 0)
                {
                    const int a = get(c, yl);
                    none = a < 0;
                    qlo = a;
                }
                if (yr < n)
                {
                    const int a = get(c, yr);
                    none = none || a < 0;
                    qhi = a;
                }
                if (!none)
                    for (int q = qlo + sub; q <= qhi; q += G)
                    {
                        const int v = cost(c, mid, q);
                        if (v < bc)
                        {
                            bc = v;
                            bq = q;
                        }
                    }
            }
            group_min(bc, bq, G);
            if (active && sub == 0)
                put(c, mid, bq, bc);
        }
        __syncthreads(); // this level's rows are read by the next
    }
}


template <bool kB, bool kM>
__global__ void __launch_bounds__(kBlock) edt_pass_y(vc_work W)
{
    extern __shared__ int16_t colz[]; // [ny][SY], SY >= TZ
    const uint32_t per = (W.total + 7) / 8;
    for (uint32_t L = blockIdx.x; L < 8 * per; L += gridDim.x)
    {
        const uint32_t b = vc_xcd_item(L, per);
        if (b >= W.total)
            continue;
        int vi;
        uint32_t T;
        vc_item<kB>(W, b, vi, T);
        const vc_vdesc d = vc_desc<kB>(W, vi);
        const int ny = d.ny, nz = d.nz, TZ = d.tzy;
        const int SY = vc_edt_ystride(TZ); // padded LDS row
        const int nzt = (nz + TZ - 1) / TZ;
        const int z0 = (int)(T % nzt) * TZ;
        const int x = (int)(T / nzt);
        const int64_t base = (int64_t)x * ny * nz;
        const int16_t* zq = reinterpret_cast<const int16_t*>(d.edt);
        {
            // staging: TZ-wide coalesced row segments, 16 independent loads
            // in flight per thread before the LDS stores
            const int tz = threadIdx.x % TZ;
            const int ty0 = threadIdx.x / TZ;
            const int tys = blockDim.x / TZ;
            const int z = z0 + tz;
            const bool zin = z < nz;
            constexpr int U = 16;
            for (int y = ty0; y < ny; y += tys * U)
            {
                int16_t v[U];
#pragma unroll
                for (int u = 0; u < U; ++u)
                {
                    const int yy = y + u * tys;
                    v[u] = (zin && yy < ny) ? zq[base + (int64_t)yy * nz + z]
                                            : (int16_t)-1;
                }
#pragma unroll
                for (int u = 0; u < U; ++u)
                {
                    const int yy = y + u * tys;
                    int16_t e = v[u];
                    if (kM)
                        e = v[u] >= 0 ? (int16_t)(z - v[u]) : (int16_t)-32768;
                    if (yy < ny)
                        colz[yy * SY + tz] = e;
                }
            }
        }
        __syncthreads();
        // x-pass tile of voxel (y, z): flagged when its (y, z) cost exceeds
        // kDeepX^2 or it has none
        const int nztx = (nz + d.tzx - 1) / d.tzx;
        auto flag_x = [&](int y, int z, bool none, int cost) {
            if (none || cost > kDeepX * kDeepX)
                d.fx[(int64_t)y * nztx + z / d.tzx] = 1;
        };
        if (d.fy[T])
        {
            // a thick tile: every row of its columns by divide and conquer
            const int ncol = min(TZ, nz - z0);
            uint32_t* ftc = d.ft + base + z0;
            auto get = [&](int c, int y) -> int {
                const uint32_t v = ftc[(int64_t)y * nz + c];
                return v == VC_NONE ? -1 : (int)(v >> 16);
            };
            if (kM)
                column_argmins(
                    ny, ncol, TZ, kNoBg,
                    [&](int c, int y, int q) -> int {
                        const int e = colz[q * SY + c];
                        return __mul24(e, e) + (y - q) * (y - q);
                    },
                    [&](int c, int y, int q, int v) {
                        ftc[(int64_t)y * nz + c] =
                            q < 0 ? VC_NONE
                                  : ((uint32_t)q << 16) |
                                        (uint32_t)(uint16_t)(z0 + c - colz[q * SY + c]);
                        flag_x(y, z0 + c, q < 0, v);
                    },
                    get);
            else
                column_argmins(
                    ny, ncol, TZ, kInf,
                    [&](int c, int y, int q) -> int {
                        const int e = colz[q * SY + c];
                        return e < 0 ? kInf
                                     : (z0 + c - e) * (z0 + c - e) + (y - q) * (y - q);
                    },
                    [&](int c, int y, int q, int v) {
                        ftc[(int64_t)y * nz + c] =
                            q < 0 ? VC_NONE
                                  : ((uint32_t)q << 16) | (uint32_t)(uint16_t)colz[q * SY + c];
                        flag_x(y, z0 + c, q < 0, v);
                    },
                    get);
            if (threadIdx.x == 0)
                d.fy[T] = 0;
            continue; // column_argmins ended with a barrier
        }
        // search lanes: each wavefront takes blocks of ZG z x YG y (8 x 8
        // for TZ >= 8); a compact block keeps the lanes' search lengths
        // close, so fewer lanes idle while the longest search runs
        const int lzg = min(__builtin_ctz((unsigned)TZ), 3);
        const int ZG = 1 << lzg, YG = 64 >> lzg;
        const int lngz = __builtin_ctz((unsigned)TZ) - lzg; // z blocks: 2^lngz
        const int nblk = ((ny + YG - 1) / YG) << lngz;
        const int lane = threadIdx.x & 63;
        for (int B = (int)(threadIdx.x >> 6); B < nblk; B += (int)(blockDim.x >> 6))
        {
            const int y = (B >> lngz) * YG + (lane >> lzg);
            const int tz = ((B & ((1 << lngz) - 1)) << lzg) + (lane & (ZG - 1));
            const int z = z0 + tz;
            if (y >= ny || z >= nz)
                continue;
            if (kM)
            {
                int best = kNoBg;
                int by = -1;
                {
                    const int e = colz[y * SY + tz];
                    const int c = __mul24(e, e);
                    if (c < kNoBg)
                    {
                        best = c;
                        by = y;
                    }
                }
                // distance dd: lower index first (ties keep the lowest y)
                auto step = [&](int dd) {
                    const int d2 = dd * dd;
                    const int yl = y - dd, yh = y + dd;
                    const int el = colz[max(yl, 0) * SY + tz];
                    const int eh = colz[min(yh, ny - 1) * SY + tz];
                    const int vl = yl >= 0 ? __mul24(el, el) + d2 : kNoBg;
                    const int vh = yh < ny ? __mul24(eh, eh) + d2 : kNoBg;
                    const bool tl = vl <= best && vl < kNoBg;
                    best = tl ? vl : best;
                    by = tl ? yl : by;
                    const bool th = vh < best;
                    best = th ? vh : best;
                    by = th ? yh : by;
                };
                // two distances per trip: every candidate at distance dd
                // costs >= dd^2, so one evaluated past the stopping distance
                // (or past the column ends, where both sides are out of
                // range) never wins
                const int ymax = max(y, ny - 1 - y);
                for (int dd = 1; dd <= ymax; dd += 2)
                {
                    if (dd * dd > best)
                        break;
                    step(dd);
                    step(dd + 1);
                }
                uint32_t outv = VC_NONE;
                if (by >= 0)
                    outv = ((uint32_t)by << 16) |
                           (uint32_t)(uint16_t)(z - colz[by * SY + tz]);
                d.ft[base + (int64_t)y * nz + z] = outv;
                flag_x(y, z, by < 0, best);
            }
            else
            {
                int best = kInf;
                int by = -1;
                {
                    const int q = colz[y * SY + tz];
                    if (q >= 0)
                    {
                        best = (z - q) * (z - q);
                        by = y;
                    }
                }
                // branch-free search body (clamped loads, selects): lower
                // index first, so ties keep the lowest y
                const int ymax = max(y, ny - 1 - y);
                for (int dd = 1; dd <= ymax; ++dd)
                {
                    const int d2 = dd * dd;
                    if (d2 > best)
                        break;
                    const int yl = y - dd, yh = y + dd;
                    const int ql = colz[max(yl, 0) * SY + tz];
                    const int qh = colz[min(yh, ny - 1) * SY + tz];
                    const int vl =
                        (yl >= 0 && ql >= 0) ? (z - ql) * (z - ql) + d2 : kInf;
                    const int vh =
                        (yh < ny && qh >= 0) ? (z - qh) * (z - qh) + d2 : kInf;
                    const bool tl = vl <= best && vl < kInf;
                    best = tl ? vl : best;
                    by = tl ? yl : by;
                    const bool th = vh < best;
                    best = th ? vh : best;
                    by = th ? yh : by;
                }
                uint32_t outv = VC_NONE;
                if (by >= 0)
                    outv = ((uint32_t)by << 16) |
                           (uint32_t)(uint16_t)colz[by * SY + tz];
                d.ft[base + (int64_t)y * nz + z] = outv;
                flag_x(y, z, by < 0, best);
            }
        }
        __syncthreads(); // colz is rewritten by the next item
    }
}

// columns along x: item = TZ consecutive z at fixed y, all x.  Rewrites ft in
// place (the workgroup stages exactly the elements it later overwrites) and
// overwrites edt once every z' of the volume has been consumed by pass y.
//
// The staged element is the offset pair (y - y', z - z') as two int16 (axes
// are at most VC_MAX_AXIS = 32767 long), so a candidate's cost dx^2 + dy^2 +
// dz^2 is one saturating dot product, sdot2(p, p, dx^2).  A column without
// background is stored as (-32768, -32768): its cost 2^31 + dx^2 saturates
// to kInf.  Real costs are squared grid distances, below the squared
// diagonal (<= INT32_MAX, checked at volume creation; INT32_MAX itself is
// 7 mod 8, not a sum of three squares), so they never reach kInf.
static_assert(VC_MAX_AXIS <= 32767, "offset pairs must fit int16");
typedef short vc_s16x2 __attribute__((ext_vector_type(2)));
constexpr uint32_t kNoFeat = 0x80008000u;

__device__ __forceinline__ int dot_cost(uint32_t p, int d2)
{
    const vc_s16x2 a = __builtin_bit_cast(vc_s16x2, p);
    return __builtin_amdgcn_sdot2(a, a, d2, true);
}

template <bool kB>
__global__ void __launch_bounds__(kBlock) edt_pass_x(vc_work W)
{
    extern __shared__ uint32_t colyz[]; // [nx][TZ]
    const uint32_t per = (W.total + 7) / 8;
    for (uint32_t L = blockIdx.x; L < 8 * per; L += gridDim.x)
    {
        const uint32_t b = vc_xcd_item(L, per);
        if (b >= W.total)
            continue;
        int vi;
        uint32_t T;
        vc_item<kB>(W, b, vi, T);
        const vc_vdesc d = vc_desc<kB>(W, vi);
        const int nx = d.nx, ny = d.ny, nz = d.nz, TZ = d.tzx;
        const int nzt = (nz + TZ - 1) / TZ;
        const int z0 = (int)(T % nzt) * TZ;
        const int y = (int)(T / nzt);
        const int tz = threadIdx.x % TZ;
        const int tx0 = threadIdx.x / TZ;
        const int txs = blockDim.x / TZ;
        const int z = z0 + tz;
        const bool zin = z < nz;
        const int64_t sx = (int64_t)ny * nz;
        const int64_t off = (int64_t)y * nz + z;
        uint32_t* ft = d.ft;

        constexpr int U = 16;
        for (int x = tx0; x < nx; x += txs * U)
        {
            uint32_t v[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                const int xx = x + u * txs;
                v[u] = (zin && xx < nx) ? ft[(int64_t)xx * sx + off] : VC_NONE;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                const int xx = x + u * txs;
                const uint32_t e =
                    v[u] == VC_NONE
                        ? kNoFeat
                        : ((uint32_t)(uint16_t)(y - (int)(v[u] >> 16)) << 16) |
                              (uint32_t)(uint16_t)(z - (int)(v[u] & 0xFFFFu));
                if (xx < nx)
                    colyz[xx * TZ + tz] = e;
            }
        }
        __syncthreads();
        if (d.fx[T])
        {
            // a thick tile: every row of its columns by divide and conquer
            // from the staged pairs (a row's A comes back from its written
            // feature)
            const int ncol = min(TZ, nz - z0);
            const int sh = d.pk.by + d.pk.bz;
            uint32_t* ftc = ft + (int64_t)y * nz + z0;
            column_argmins(
                nx, ncol, TZ, kInf,
                [&](int c, int xx, int q) -> int {
                    return dot_cost(colyz[q * TZ + c], (xx - q) * (xx - q));
                },
                [&](int c, int xx, int q, int v) {
                    const int64_t o = (int64_t)xx * sx + (int64_t)y * nz + z0 + c;
                    if (q < 0)
                    {
                        ft[o] = VC_NONE;
                        d.edt[o] = -1;
                    }
                    else
                    {
                        const uint32_t e = colyz[q * TZ + c];
                        ft[o] = vc_pack_xyz(d.pk, q, y - (int)(int16_t)(e >> 16),
                                            z0 + c - (int)(int16_t)(e & 0xFFFFu));
                        d.edt[o] = (int32_t)v;
                    }
                },
                [&](int c, int xx) -> int {
                    const uint32_t v = ftc[(int64_t)xx * sx + c];
                    return v == VC_NONE ? -1 : (int)(v >> sh);
                });
            if (threadIdx.x == 0)
                d.fx[T] = 0;
            continue; // column_argmins ended with a barrier
        }
        if (zin)
            for (int x = tx0; x < nx; x += txs)
            {
                int best = kInf;
                int bxv = -1;
                // cost of candidate xx plus d2 (kInf without background)
                auto cost = [&](int xx, int d2) -> int {
                    return dot_cost(colyz[xx * TZ + tz], d2);
                };
                {
                    const int c = cost(x, 0);
                    if (c < kInf)
                    {
                        best = c;
                        bxv = x;
                    }
                }
                // branch-free search body (clamped loads, selects): lower
                // index first, so ties keep the lowest x
                auto step = [&](int dd) {
                    const int d2 = dd * dd;
                    const int xl = x - dd, xh = x + dd;
                    const int cl = cost(max(xl, 0), d2), ch = cost(min(xh, nx - 1), d2);
                    const int vl = (xl >= 0) ? cl : kInf;
                    const int vh = (xh < nx) ? ch : kInf;
                    const bool tl = vl <= best && vl < kInf;
                    best = tl ? vl : best;
                    bxv = tl ? xl : bxv;
                    const bool th = vh < best;
                    best = th ? vh : best;
                    bxv = th ? xh : bxv;
                };
                // two distances per trip, as in the y pass
                const int xmax = max(x, nx - 1 - x);
                for (int dd = 1; dd <= xmax; dd += 2)
                {
                    if (dd * dd > best)
                        break;
                    step(dd);
                    step(dd + 1);
                }
                const int64_t o = (int64_t)x * sx + off;
                if (bxv < 0)
                {
                    ft[o] = VC_NONE;
                    d.edt[o] = -1;
                }
                else
                {
                    const uint32_t v = colyz[bxv * TZ + tz];
                    const int yy = y - (int)(int16_t)(v >> 16);
                    const int zz = z - (int)(int16_t)(v & 0xFFFFu);
                    ft[o] = vc_pack_xyz(d.pk, bxv, yy, zz);
                    d.edt[o] = (int32_t)best;
                }
            }
        __syncthreads(); // colyz is rewritten by the next item
    }
}

// Per-voxel measure, one wavefront per strip of kMeasZ = 62 z x kMeasY y x
// kMeasXC x: lanes hold z0 - 1 .. z0 + 62 (lanes 0 and 63 are the z halo),
// registers hold slices x - 1, x, x + 1 (and x + 2 in flight) of the
// kMeasY + 2 rows y0 - 1 .. y0 + kMeasY, and z neighbours come from the
// adjacent lanes: no LDS and no barrier.  Loads are unconditional from clamped, in-bounds addresses and
// the registers keep the raw values; out-of-grid neighbours become kOob
// where they are used (a load under a branch, or a select right after it,
// makes the compiler wait for every outstanding load).  A voxel q is
// background iff FT(q) == q, so no occupancy is read.
constexpr int kMeasR = kMeasY + 2; // rows held: the strip's and y -+ 1
constexpr int kMeasWaves = 4; // independent strips per workgroup

template <bool kB>
__global__ void __launch_bounds__(64 * kMeasWaves) measure_lam2(vc_work W)
{
    const int wave = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    const uint32_t per = (W.total + 7) / 8;
    // slot of wave w of workgroup g: g + w * gridDim.x (+ rounds), which
    // keeps g's XCD item group (gridDim.x is a multiple of 8)
    for (uint32_t L = blockIdx.x + gridDim.x * (uint32_t)wave; L < 8 * per;
         L += kMeasWaves * gridDim.x)
    {
        const uint32_t b = vc_xcd_item(L, per);
        if (b >= W.total)
            continue;
        int vi;
        uint32_t T;
        vc_item<kB>(W, b, vi, T);
        const vc_vdesc d = vc_desc<kB>(W, vi);
        const int nx = d.nx, ny = d.ny, nz = d.nz;
        const vc_pack pk = d.pk;
        const uint32_t* ft = d.ft;
        const int nzt = (nz + kMeasZ - 1) / kMeasZ;
        const int nyt = (ny + kMeasY - 1) / kMeasY;
        const int z0 = (int)(T % nzt) * kMeasZ;
        const int y0 = (int)((T / nzt) % nyt) * kMeasY;
        const int xa = (int)(T / ((uint32_t)nzt * nyt)) * kMeasXC;
        const int xb = min(nx, xa + kMeasXC);
        const int z = z0 - 1 + lane;
        const bool zok = lane >= 1 && lane <= kMeasZ && z < nz; // computed lanes
        // element indices fit 32 bits (N < 2^31)
        const uint32_t sx = (uint32_t)ny * (uint32_t)nz;
        const int zc = min(max(z, 0), nz - 1);
        // row i is y0 - 1 + i; slices past xb (whose next neighbour is read)
        // clamp to it, so prefetches beyond the chunk fetch nothing new
        uint32_t roff[kMeasR];
#pragma unroll
        for (int i = 0; i < kMeasR; ++i)
            roff[i] = (uint32_t)min(max(y0 - 1 + i, 0), ny - 1) * (uint32_t)nz + (uint32_t)zc;
        const int xlast = min(xb, nx - 1);
        auto ld = [&](int xx, int i) -> uint32_t {
            return ft[(uint32_t)min(max(xx, 0), xlast) * sx + roff[i]];
        };
        // a voxel's feature, unpacked, and whether the voxel is inside (in
        // the grid, with a feature, not background: FT(v) != v).  A pair of
        // 6-neighbours counts for both voxels' measure iff both are inside,
        // so each pair is priced once, at its lower voxel (+x, +y, +z), and
        // reaches the upper voxel from the previous slice, the row below or
        // the lane below.
        struct Ft
        {
            int x, y, z;
            bool in;
        };
        auto feat = [&](uint32_t g, uint32_t self, bool grid) -> Ft {
            Ft f;
            vc_unpack_xyz(pk, g, f.x, f.y, f.z);
            f.in = grid && g != VC_NONE && g != self;
            return f;
        };
        auto dist = [&](const Ft& a, const Ft& b) -> uint32_t {
            const uint32_t dd = (uint32_t)((a.x - b.x) * (a.x - b.x)) +
                                (uint32_t)((a.y - b.y) * (a.y - b.y)) +
                                (uint32_t)((a.z - b.z) * (a.z - b.z));
            return (a.in && b.in) ? dd : 0u;
        };
        const bool zin = z >= 0 && z < nz;
        bool rin[kMeasR];
#pragma unroll
        for (int i = 0; i < kMeasR; ++i)
            rin[i] = (unsigned)(y0 - 1 + i) < (unsigned)ny;
        const uint32_t sxp = 1u << (pk.by + pk.bz), syp = 1u << pk.bz;
        // packed index of (x, y0 - 1 + r, z) = pack(x, y0, zc) + (r - 1) syp
        // (only compared where the voxel is in the grid)
        const uint32_t me0 = vc_pack_xyz(pk, xa, y0, zc);
        uint32_t C[kMeasR], Q[kMeasR]; // raw slices x, x + 1
        uint32_t pxb[kMeasR];          // pair (x - 1, x) of each strip row
        {
            uint32_t Pv[kMeasR];
#pragma unroll
            for (int i = 0; i < kMeasR; ++i)
            {
                Pv[i] = ld(xa - 1, i);
                C[i] = ld(xa, i);
                Q[i] = ld(xa + 1, i);
            }
#pragma unroll
            for (int r = 1; r <= kMeasY; ++r)
            {
                const uint32_t self = me0 + (uint32_t)(r - 1) * syp;
                pxb[r] = dist(feat(Pv[r], self - sxp, xa > 0 && rin[r] && zin),
                              feat(C[r], self, rin[r] && zin));
            }
        }
        for (int x = xa; x < xb; ++x)
        {
            uint32_t F[kMeasR];
#pragma unroll
            for (int i = 0; i < kMeasR; ++i)
                F[i] = ld(x + 2, i);
            const uint32_t meX = me0 + (uint32_t)(x - xa) * sxp;
            Ft c[kMeasR];
#pragma unroll
            for (int r = 0; r < kMeasR; ++r)
                c[r] = feat(C[r], meX + (uint32_t)(r - 1) * syp, rin[r] && zin);
            uint32_t py[kMeasR - 1]; // pair (row r, row r + 1)
#pragma unroll
            for (int r = 0; r + 1 < kMeasR; ++r)
                py[r] = dist(c[r], c[r + 1]);
            const bool xnext = x + 1 < nx;
#pragma unroll
            for (int r = 1; r <= kMeasY; ++r)
            {
                const int y = y0 - 1 + r;
                // pair (z, z + 1) with the lane above, (z - 1, z) from the
                // lane below (all lanes shuffle)
                Ft zn;
                vc_unpack_xyz(pk, __shfl_down(C[r], 1), zn.x, zn.y, zn.z);
                zn.in = __shfl_down((int)c[r].in, 1) != 0;
                const uint32_t pz = dist(c[r], zn);
                const uint32_t pzb = __shfl_up(pz, 1);
                const uint32_t self = meX + (uint32_t)(r - 1) * syp;
                const uint32_t px = dist(c[r], feat(Q[r], self + sxp, xnext && rin[r] && zin));
                const uint32_t m =
                    max(max(px, pxb[r]), max(max(py[r - 1], py[r]), max(pz, pzb)));
                pxb[r] = px;
                if (zok && y < ny)
                    d.lam2[(uint32_t)x * sx + roff[r]] = m;
            }
#pragma unroll
            for (int i = 0; i < kMeasR; ++i)
            {
                C[i] = Q[i];
                Q[i] = F[i];
            }
        }
    }
}

__global__ void __launch_bounds__(kBlock)
    ft_linear(const uint32_t* __restrict__ ft, int32_t* __restrict__ out,
              int ny, int nz, int64_t N, vc_pack pk)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const uint32_t f = ft[i];
    if (f == VC_NONE)
    {
        out[i] = -1;
        return;
    }
    int x, y, z;
    vc_unpack_xyz(pk, f, x, y, z);
    out[i] = (int32_t)(((int64_t)x * ny + y) * nz + z);
}

} // namespace

namespace
{
__global__ void __launch_bounds__(256) fill_u8(uint8_t* p, int64_t n, uint8_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
} // namespace

int vc_fill_u8(vc_ctx* ctx, uint8_t* p, size_t n, uint8_t v)
{
    if (n == 0)
        return VC_OK;
    const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    fill_u8<<<g, 256, 0, ctx->stream>>>(p, (int64_t)n, v);
    VC_CHECK_LAUNCH(ctx);
    return VC_OK;
}

int vc_edt_sched(vc_ctx* ctx, vc_sched* s)
{
    VC_HIP(ctx, hipSetDevice(ctx->device));
    // No temporaries: pass z writes int16 z' into the first half of each
    // edt buffer, pass y writes packed (y', z') into the ft buffer, and pass
    // x rewrites ft in place and then edt.
    const int pc = s->sc.edt_grid_per_cu;
    unsigned g = vc_sched_grid(ctx, s, VK_EZ, pc);
    if (g)
    {
        if (s->n > 1)
            edt_pass_z<true><<<g, kBlock, s->lds[VK_EZ], ctx->stream>>>(
                vc_sched_work(s, VK_EZ), s->ez_stride);
        else
            edt_pass_z<false><<<g, kBlock, s->lds[VK_EZ], ctx->stream>>>(
                vc_sched_work(s, VK_EZ), s->ez_stride);
        VC_CHECK_LAUNCH(ctx);
    }
    g = vc_sched_grid(ctx, s, VK_EY, pc);
    if (g)
    {
        bool m24 = true; // signed offsets and their squares fit
        for (const vc_vol* v : s->vols)
            m24 = m24 && v->ny <= 16384 && v->nz <= 16384;
        const size_t ly = s->lds[VK_EY];
        const vc_work wy = vc_sched_work(s, VK_EY);
        if (s->n > 1 && m24)
            edt_pass_y<true, true><<<g, kBlock, ly, ctx->stream>>>(wy);
        else if (s->n > 1)
            edt_pass_y<true, false><<<g, kBlock, ly, ctx->stream>>>(wy);
        else if (m24)
            edt_pass_y<false, true><<<g, kBlock, ly, ctx->stream>>>(wy);
        else
            edt_pass_y<false, false><<<g, kBlock, ly, ctx->stream>>>(wy);
        VC_CHECK_LAUNCH(ctx);
    }
    g = vc_sched_grid(ctx, s, VK_EX, pc);
    if (g)
    {
        if (s->n > 1)
            edt_pass_x<true><<<g, kBlock, s->lds[VK_EX], ctx->stream>>>(
                vc_sched_work(s, VK_EX));
        else
            edt_pass_x<false><<<g, kBlock, s->lds[VK_EX], ctx->stream>>>(
                vc_sched_work(s, VK_EX));
        VC_CHECK_LAUNCH(ctx);
    }
    for (vc_vol* v : s->vols)
        v->edt_done = true;
    return VC_OK;
}

int vc_measure_sched(vc_ctx* ctx, vc_sched* s)
{
    VC_HIP(ctx, hipSetDevice(ctx->device));
    // the scheduler counts one slot per strip; a workgroup runs kMeasWaves
    // strips (grid a multiple of 8 for the XCD mapping)
    const unsigned gs = vc_sched_grid(ctx, s, VK_MS, s->sc.edt_grid_per_cu);
    const unsigned g = 8 * ((gs + 8 * kMeasWaves - 1) / (8 * kMeasWaves));
    if (gs)
    {
        if (s->n > 1)
            measure_lam2<true><<<g, 64 * kMeasWaves, 0, ctx->stream>>>(
                vc_sched_work(s, VK_MS));
        else
            measure_lam2<false><<<g, 64 * kMeasWaves, 0, ctx->stream>>>(
                vc_sched_work(s, VK_MS));
        VC_CHECK_LAUNCH(ctx);
    }
    for (vc_vol* v : s->vols)
        v->measure_kind = 1;
    return vc_sched_sync_measure(ctx, s);
}

int vc_ft_to_linear(vc_ctx* ctx, vc_vol* v, int32_t* dst)
{
    unsigned blocks = (unsigned)((v->N + kBlock - 1) / kBlock);
    ft_linear<<<blocks, kBlock, 0, ctx->stream>>>(v->ft, dst, v->ny, v->nz,
                                                  v->N, v->pk);
    VC_CHECK_LAUNCH(ctx);
    return VC_OK;
}
