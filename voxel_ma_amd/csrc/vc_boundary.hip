// vc_boundary.hip -- boundary sites, boundary quads and the cube-complex
// Euler characteristic with the reference Surfacer's exact output order
// (src/surfacing.cpp:49-197, :323-440, :597-660), gfx950.
//
// Order.  Surfacer visits voxels x-outer / z-inner (= our linear index), the
// six face offsets -x,+x,-y,+y,-z,+z of each voxel, and for a face between
// voxels of different occupancy first the four corners m_cornersWRTNbOffset[o]
// in order, then the face; a corner or face is emitted at its first visit.
// A corner is touched by the voxels of its 2x2x2 block, and a block voxel
// touches it iff it differs from one of its three block neighbours, so the
// corner's first visitor is the first (in linear order) in-grid block voxel
// that touches it; the visit happens at that voxel's first touching offset.
// Every voxel therefore decides from its 3x3x3 neighbourhood which corners
// and faces it emits, and in which order.  Counts per 256-voxel tile, an
// exclusive scan over tiles and a per-tile scan give every output its index.
//
// Quads refer to sites through a corner-lattice map (written when the sites
// are emitted, read when the quads are).  The Euler terms follow
// computeEulerChar literally, including the (int)-truncation of the corner
// base voxel at coordinate -0.5.  C is the number of 6-components of the
// "thin inside layer" (inside voxels within Chebyshev distance 1 of an inside
// voxel with a non-inside face neighbour), by union-find.
#include <algorithm>
#include <cstring>
#include <vector>

#include "vc_internal.h"

namespace
{

constexpr int kBlock = 256;

// ---------------------------------------------------------- reference tables
// Tables of the reference Surfacer, as compile-time functions (run-time
// lookups go through packed integer constants, never through local arrays).
struct Tab
{
    int8_t off[6][3];  // face offsets o: -x, +x, -y, +y, -z, +z (include/surfacing.h:172-178)
    int8_t cwrt[6][4]; // corners shared with the neighbour across offset o (surfacing.h:185-190)
    int8_t fwrt[6];    // face index of offset o (surfacing.h:191-197)
    int8_t fcrn[6][4]; // corner slots of face fi (surfacing.h:144-149)
    int8_t sgn[8][3];  // corner c of voxel (i,j,k) is at (i,j,k) + 0.5 * sgn[c] (surfacing.h:97-119)
};
constexpr Tab kT = {
    {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}},
    {{2, 3, 7, 6}, {0, 1, 5, 4}, {4, 6, 2, 0}, {1, 3, 7, 5}, {0, 2, 3, 1}, {5, 7, 6, 4}},
    {3, 2, 4, 5, 0, 1},
    {{0, 1, 3, 2}, {4, 5, 7, 6}, {0, 4, 5, 1}, {2, 6, 7, 3}, {4, 0, 2, 6}, {5, 1, 3, 7}},
    {{1, -1, -1}, {1, 1, -1}, {-1, -1, -1}, {-1, 1, -1}, {1, -1, 1}, {1, 1, 1}, {-1, -1, 1},
     {-1, 1, 1}}};

// corner signs in closed form: x is + for slots {0,1,4,5}, y for odd slots,
// z for slots >= 4
__host__ __device__ constexpr int csx(int c) { return (c & 2) ? -1 : 1; }
__host__ __device__ constexpr int csy(int c) { return (c & 1) ? 1 : -1; }
__host__ __device__ constexpr int csz(int c) { return (c & 4) ? 1 : -1; }
constexpr bool signs_ok()
{
    for (int c = 0; c < 8; ++c)
        if (csx(c) != kT.sgn[c][0] || csy(c) != kT.sgn[c][1] || csz(c) != kT.sgn[c][2])
            return false;
    return true;
}
static_assert(signs_ok(), "corner sign formulas");

// slot ii of corner c in cwrt[o] (every corner on the side of o is listed)
constexpr int cwrt_pos(int o, int c)
{
    for (int ii = 0; ii < 4; ++ii)
        if (kT.cwrt[o][ii] == c)
            return ii;
    return -1;
}
// first-visit key 4 o + ii -> corner slot, 3 bits per key (keys 0-15, 16-23)
constexpr uint64_t key_bits(int from)
{
    uint64_t r = 0;
    for (int k = from; k < 24 && k < from + 16; ++k)
        r |= (uint64_t)kT.cwrt[k >> 2][k & 3] << (3 * (k - from));
    return r;
}
__device__ __forceinline__ int key_corner(int k)
{
    constexpr uint64_t lo = key_bits(0), hi = key_bits(16);
    return (int)(((k < 16) ? (lo >> (3 * k)) : (hi >> (3 * (k - 16)))) & 7u);
}
// the 4 corner slots of the face across offset o, 12 bits per offset
constexpr uint64_t face_bits(int from)
{
    uint64_t r = 0;
    for (int o = from; o < 6 && o < from + 5; ++o)
        for (int t = 0; t < 4; ++t)
            r |= (uint64_t)kT.fcrn[kT.fwrt[o]][t] << (12 * (o - from) + 3 * t);
    return r;
}
__device__ __forceinline__ uint32_t face_corners(int o)
{
    constexpr uint64_t lo = face_bits(0), hi = face_bits(5);
    return (uint32_t)((o < 5 ? (lo >> (12 * o)) : hi) & 0xFFFu);
}

// neighbourhood bit of offset (dx, dy, dz): dz+1 + 3(dy+1) + 9(dx+1)
__host__ __device__ constexpr int nbit(int dx, int dy, int dz)
{
    return (dz + 1) + 3 * (dy + 1) + 9 * (dx + 1);
}

struct Vox
{
    uint32_t occ;  // 27 bits, out of grid = 0
    uint32_t ing;  // 27 bits, in grid
};

__device__ __forceinline__ bool obit(uint32_t m, int dx, int dy, int dz)
{
    return (m >> nbit(dx, dy, dz)) & 1u;
}

// Decisions of one voxel as masks: new corners by first-visit key 4 o + ii
// (emitted in ascending key order; distinct corners have distinct keys) and
// new faces by offset o (emitted in ascending o).
struct Emit
{
    uint32_t keys;
    uint32_t faces;
};

__device__ __forceinline__ Emit decide(const Vox& V)
{
    Emit r{0u, 0u};
    const bool me = obit(V.occ, 0, 0, 0);
    // events of this voxel
    uint32_t ev = 0;
#pragma unroll
    for (int o = 0; o < 6; ++o)
        if (obit(V.occ, kT.off[o][0], kT.off[o][1], kT.off[o][2]) != me)
            ev |= 1u << o;
    if (!ev)
        return r;
#pragma unroll
    for (int c = 0; c < 8; ++c)
    {
        const int sx = csx(c), sy = csy(c), sz = csz(c);
        const int ox = sx > 0 ? 1 : 0, oy = sy > 0 ? 3 : 2, oz = sz > 0 ? 5 : 4;
        // key of the first touching offset
        int key;
        if (ev & (1u << ox))
            key = 4 * ox + cwrt_pos(ox, c);
        else if (ev & (1u << oy))
            key = 4 * oy + cwrt_pos(oy, c);
        else if (ev & (1u << oz))
            key = 4 * oz + cwrt_pos(oz, c);
        else
            continue;
        // an earlier (linear order = lexicographic block order) in-grid
        // block voxel touching the corner visits it first.  Block slot
        // q = 4a + 2b + e holds voxel (lx + a, ly + b, lz + e); it touches the
        // corner iff it differs from one of its three block neighbours
        // (slots q ^ 4, q ^ 2, q ^ 1); this voxel is slot q0.
        const int lx = sx > 0 ? 0 : -1, ly = sy > 0 ? 0 : -1, lz = sz > 0 ? 0 : -1;
        const int q0 = 4 * -lx + 2 * -ly + -lz;
        uint32_t m8 = 0, g8 = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q)
        {
            const int ux = lx + (q >> 2), uy = ly + ((q >> 1) & 1), uz = lz + (q & 1);
            m8 |= (uint32_t)obit(V.occ, ux, uy, uz) << q;
            g8 |= (uint32_t)obit(V.ing, ux, uy, uz) << q;
        }
        const uint32_t f4 = ((m8 >> 4) | (m8 << 4)) & 0xFFu;
        const uint32_t f2 = ((m8 >> 2) & 0x33u) | ((m8 << 2) & 0xCCu);
        const uint32_t f1 = ((m8 >> 1) & 0x55u) | ((m8 << 1) & 0xAAu);
        const uint32_t touch = (m8 ^ f4) | (m8 ^ f2) | (m8 ^ f1);
        if (!(touch & g8 & ((1u << q0) - 1u)))
            r.keys |= 1u << key;
    }
    // faces created here: +offsets, and -offsets whose neighbour is outside
#pragma unroll
    for (int o = 0; o < 6; ++o)
        if ((ev >> o) & 1u)
        {
            const bool nb_in = obit(V.ing, kT.off[o][0], kT.off[o][1], kT.off[o][2]);
            if ((o & 1) || !nb_in)
                r.faces |= 1u << o;
        }
    return r;
}

struct Dims
{
    int nx, ny, nz;
};

__device__ __forceinline__ Vox gather(const uint8_t* __restrict__ occ,
                                      const Dims& D, int x, int y, int z)
{
    Vox V{0u, 0u};
    const int64_t sy = D.nz, sx = (int64_t)D.ny * D.nz;
    const int64_t i = x * sx + y * sy + z;
#pragma unroll
    for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
            for (int dz = -1; dz <= 1; ++dz)
            {
                const bool in = (unsigned)(x + dx) < (unsigned)D.nx &&
                                (unsigned)(y + dy) < (unsigned)D.ny &&
                                (unsigned)(z + dz) < (unsigned)D.nz;
                if (in)
                {
                    V.ing |= 1u << nbit(dx, dy, dz);
                    if (occ[i + dx * sx + dy * sy + dz])
                        V.occ |= 1u << nbit(dx, dy, dz);
                }
            }
    return V;
}

// whether voxel (x, y, z) differs from one of its six face neighbours
// (outside the grid reads 0): voxels without such an event emit nothing, so
// the full 27-neighbour gather is skipped for them
__device__ __forceinline__ bool face_event(const uint8_t* __restrict__ occ,
                                           const Dims& D, int x, int y, int z)
{
    const int64_t sy = D.nz, sx = (int64_t)D.ny * D.nz;
    const int64_t i = x * sx + y * sy + z;
    const bool me = occ[i] != 0;
    const bool xm = x > 0 && occ[i - sx] != 0, xp = x + 1 < D.nx && occ[i + sx] != 0;
    const bool ym = y > 0 && occ[i - sy] != 0, yp = y + 1 < D.ny && occ[i + sy] != 0;
    const bool zm = z > 0 && occ[i - 1] != 0, zp = z + 1 < D.nz && occ[i + 1] != 0;
    return (xm != me) | (xp != me) | (ym != me) | (yp != me) | (zm != me) |
           (zp != me);
}

__device__ __forceinline__ void decode(const Dims& D, int64_t i, int& x,
                                       int& y, int& z)
{
    // N < 2^31: 32-bit division
    const uint32_t u = (uint32_t)i;
    z = (int)(u % (uint32_t)D.nz);
    const uint32_t r = u / (uint32_t)D.nz;
    y = (int)(r % (uint32_t)D.ny);
    x = (int)(r / (uint32_t)D.ny);
}

// corner lattice id of corner slot c of voxel (x, y, z)
__device__ __forceinline__ int64_t corner_id(const Dims& D, int x, int y,
                                             int z, int c)
{
    const int cx = x + (csx(c) > 0), cy = y + (csy(c) > 0), cz = z + (csz(c) > 0);
    return ((int64_t)cx * (D.ny + 1) + cy) * (D.nz + 1) + cz;
}

// block-wide exclusive scan of two counters packed as (a << 16) | b
__device__ uint32_t block_scan_excl(uint32_t v, uint32_t* sh, uint32_t* total)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1)
    {
        const uint32_t t = __shfl_up(incl, d);
        if (lane >= d)
            incl += t;
    }
    if (lane == 63)
        sh[wv] = incl;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
    {
        if (w < wv)
            off += sh[w];
        tot += sh[w];
    }
    *total = tot;
    return off + incl - v;
}

// pass 1: per-tile totals of (sites, faces)
__global__ void __launch_bounds__(kBlock)
    bnd_count(const uint8_t* __restrict__ occ, Dims D, int64_t N,
              uint32_t* __restrict__ tile_sites, uint32_t* __restrict__ tile_faces)
{
    __shared__ uint32_t sh[kBlock / 64];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t ns = 0, nf = 0;
    if (i < N)
    {
        int x, y, z;
        decode(D, i, x, y, z);
        if (face_event(occ, D, x, y, z))
        {
            const Emit e = decide(gather(occ, D, x, y, z));
            ns = (uint32_t)__builtin_popcount(e.keys);
            nf = (uint32_t)__builtin_popcount(e.faces);
        }
    }
    uint32_t tot;
    block_scan_excl((ns << 16) | nf, sh, &tot);
    if (threadIdx.x == 0)
    {
        tile_sites[blockIdx.x] = tot >> 16;
        tile_faces[blockIdx.x] = tot & 0xFFFFu;
    }
}

// ---- vectorised passes (nz a multiple of 8): one thread per 8 consecutive
// z of a row, in linear order.  The 3 x 3 neighbouring rows are read as one
// 8-byte load each plus the two edge bytes, packed into 10-bit windows (bit
// b = occupancy of z0 - 1 + b, outside the grid = 0).
constexpr int kVz = 8;

// The thread of chunk c + 1 is the next lane (threads are chunks in linear
// order), and chunk c + 1 is the next 8 z of the same row unless c ends its
// row: the two edge bits come from the neighbouring lanes' middle bytes, and
// only lanes 0 and 63 load them.  Every lane of the wavefront calls it (the
// shuffles); `live` is false for threads past the last chunk, which load
// nothing (they are the highest lanes, so no live lane needs their bits: a
// live chunk whose successor is past the end ends its row).
__device__ __forceinline__ uint32_t row_win(const uint8_t* __restrict__ occ,
                                            const Dims& D, int x, int y, int z0,
                                            bool live)
{
    const bool in = live && (unsigned)x < (unsigned)D.nx && (unsigned)y < (unsigned)D.ny;
    const uint8_t* r = occ + (in ? ((int64_t)x * D.ny + y) * D.nz : 0);
    const uint64_t mid = in ? *reinterpret_cast<const uint64_t*>(r + z0) : 0ull;
    // one bit per non-zero byte, gathered into the top byte
    uint64_t t = mid | (mid >> 4);
    t |= t >> 2;
    t |= t >> 1;
    t &= 0x0101010101010101ull;
    const uint32_t m8 = (uint32_t)((t * 0x0102040810204080ull) >> 56);
    const int lane = threadIdx.x & 63;
    const uint32_t prev = __shfl_up(m8, 1u);
    const uint32_t next = __shfl_down(m8, 1u);
    uint32_t lo = 0, hi = 0;
    if (in && z0 > 0)
        lo = lane > 0 ? (prev >> 7) & 1u : (uint32_t)(r[z0 - 1] != 0);
    if (in && z0 + kVz < D.nz)
        hi = lane < 63 ? next & 1u : (uint32_t)(r[z0 + kVz] != 0);
    return lo | (m8 << 1) | (hi << 9);
}

struct Chunk
{
    uint32_t w[3][3]; // windows of rows (x + dx - 1, y + dy - 1)
    int x, y, z0;
};

// every lane of the wavefront calls it (row_win); false past the last chunk
__device__ __forceinline__ bool load_chunk(const uint8_t* __restrict__ occ,
                                           const Dims& D, int64_t c, Chunk& K)
{
    const int nzc = D.nz / kVz;
    const bool live = c < (int64_t)D.nx * D.ny * nzc;
    const int64_t row = (live ? c : 0) / nzc;
    K.z0 = (int)((live ? c : 0) - row * nzc) * kVz;
    K.x = (int)(row / D.ny);
    K.y = (int)(row - (int64_t)K.x * D.ny);
#pragma unroll
    for (int dx = 0; dx < 3; ++dx)
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
            K.w[dx][dy] = row_win(occ, D, K.x + dx - 1, K.y + dy - 1, K.z0, live);
    return live;
}

// 8 x 8 bit transpose: bit 8 r + c of x becomes bit 8 c + r
__device__ __forceinline__ uint64_t transpose8(uint64_t x)
{
    uint64_t t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
    x ^= t ^ (t << 7);
    t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
    x ^= t ^ (t << 14);
    t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
    x ^= t ^ (t << 28);
    return x;
}

// The decisions of `decide` for the 8 voxels of a chunk at once, bit-sliced:
// every quantity is an 8-bit word whose bit j belongs to voxel z0 + j, so
// voxels with events cost the wavefront no divergent work.  em[j] gets
// voxel j's corner keys (bits 0..23) and face offsets (bits 24..29); returns
// the chunk's (sites << 16) | faces.
__device__ __forceinline__ uint32_t decide_chunk(const Chunk& K, const Dims& D,
                                                 uint32_t em[kVz])
{
    // occupancy of offset (dx, dy, dz) per voxel, and its in-grid mask
    auto O = [&](int dx, int dy, int dz) -> uint32_t {
        return (K.w[dx + 1][dy + 1] >> (1 + dz)) & 0xFFu;
    };
    const uint32_t zlo = K.z0 == 0 ? 0xFEu : 0xFFu;          // z - 1 in grid
    const uint32_t zhi = K.z0 + kVz >= D.nz ? 0x7Fu : 0xFFu; // z + 1 in grid
    auto G = [&](int dx, int dy, int dz) -> uint32_t {
        const bool rin = (unsigned)(K.x + dx) < (unsigned)D.nx &&
                         (unsigned)(K.y + dy) < (unsigned)D.ny;
        return rin ? (dz < 0 ? zlo : (dz > 0 ? zhi : 0xFFu)) : 0u;
    };
    const uint32_t me = O(0, 0, 0);
    uint32_t ev[6];
#pragma unroll
    for (int o = 0; o < 6; ++o)
        ev[o] = me ^ O(kT.off[o][0], kT.off[o][1], kT.off[o][2]);
    uint32_t key[24];
#pragma unroll
    for (int k = 0; k < 24; ++k)
        key[k] = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c)
    {
        const int sx = csx(c), sy = csy(c), sz = csz(c);
        const int ox = sx > 0 ? 1 : 0, oy = sy > 0 ? 3 : 2, oz = sz > 0 ? 5 : 4;
        const int lx = sx > 0 ? 0 : -1, ly = sy > 0 ? 0 : -1, lz = sz > 0 ? 0 : -1;
        const int q0 = 4 * -lx + 2 * -ly + -lz;
        uint32_t W[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
            W[q] = O(lx + (q >> 2), ly + ((q >> 1) & 1), lz + (q & 1));
        // no earlier in-grid block voxel touches the corner
        uint32_t first = 0xFFu;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (q < q0)
            {
                const uint32_t touch = (W[q] ^ W[q ^ 4]) | (W[q] ^ W[q ^ 2]) |
                                       (W[q] ^ W[q ^ 1]);
                first &= ~(touch & G(lx + (q >> 2), ly + ((q >> 1) & 1), lz + (q & 1)));
            }
        // keyed at the first touching offset among ox, oy, oz
        const uint32_t ax = first & ev[ox];
        const uint32_t ay = first & ~ev[ox] & ev[oy];
        const uint32_t az = first & ~ev[ox] & ~ev[oy] & ev[oz];
        key[4 * ox + cwrt_pos(ox, c)] |= ax;
        key[4 * oy + cwrt_pos(oy, c)] |= ay;
        key[4 * oz + cwrt_pos(oz, c)] |= az;
    }
    // faces: +offsets, and -offsets whose neighbour is outside the grid
    uint64_t F = 0;
#pragma unroll
    for (int o = 0; o < 6; ++o)
    {
        const uint32_t f = (o & 1) ? ev[o]
                                   : ev[o] & ~G(kT.off[o][0], kT.off[o][1], kT.off[o][2]);
        F |= (uint64_t)(f & 0xFFu) << (8 * o);
    }
    uint64_t M[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 24; ++k)
        M[k >> 3] |= (uint64_t)(key[k] & 0xFFu) << (8 * (k & 7));
    const uint32_t ns = (uint32_t)(__popcll(M[0]) + __popcll(M[1]) + __popcll(M[2]));
    const uint32_t nf = (uint32_t)__popcll(F);
    if (ns | nf)
    {
        const uint64_t T0 = transpose8(M[0]), T1 = transpose8(M[1]),
                       T2 = transpose8(M[2]), T3 = transpose8(F);
#pragma unroll
        for (int j = 0; j < kVz; ++j)
            em[j] = (uint32_t)((T0 >> (8 * j)) & 0xFFu) |
                    (uint32_t)((T1 >> (8 * j)) & 0xFFu) << 8 |
                    (uint32_t)((T2 >> (8 * j)) & 0xFFu) << 16 |
                    (uint32_t)((T3 >> (8 * j)) & 0xFFu) << 24;
    }
    return (ns << 16) | nf;
}

// emission of one voxel's sites (coordinates + corner map) and quads
__device__ __forceinline__ void put_sites(const Emit& e, const Dims& D, int x, int y,
                                          int z, uint32_t s, float* __restrict__ sites,
                                          uint32_t* __restrict__ cmap)
{
    for (uint32_t k = e.keys; k; k &= k - 1, ++s)
    {
        const int cc = key_corner(__builtin_ctz(k));
        float* p = sites + 3 * (int64_t)s;
        p[0] = (float)x + 0.5f * (float)csx(cc);
        p[1] = (float)y + 0.5f * (float)csy(cc);
        p[2] = (float)z + 0.5f * (float)csz(cc);
        cmap[corner_id(D, x, y, z, cc)] = s;
    }
}

__device__ __forceinline__ void put_quads(const Emit& e, const Dims& D, int x, int y,
                                          int z, uint32_t f,
                                          const uint32_t* __restrict__ cmap,
                                          int32_t* __restrict__ quads)
{
    for (uint32_t k = e.faces; k; k &= k - 1, ++f)
    {
        const uint32_t fc = face_corners(__builtin_ctz(k));
        int4 q;
        q.x = (int32_t)cmap[corner_id(D, x, y, z, fc & 7u)];
        q.y = (int32_t)cmap[corner_id(D, x, y, z, (fc >> 3) & 7u)];
        q.z = (int32_t)cmap[corner_id(D, x, y, z, (fc >> 6) & 7u)];
        q.w = (int32_t)cmap[corner_id(D, x, y, z, (fc >> 9) & 7u)];
        *reinterpret_cast<int4*>(quads + 4 * (int64_t)f) = q;
    }
}

// pass 1 (vectorised): per-chunk counts (sites << 16 | faces), per-tile
// totals, and for chunks that emit anything each voxel's decision masks
// (corner keys in bits 0..23, face offsets in bits 24..29; 0 without an
// event), so the emit passes only scatter; chunks without output are skipped
__global__ void __launch_bounds__(kBlock)
    bnd_count_v(const uint8_t* __restrict__ occ, Dims D,
                uint32_t* __restrict__ chunk_cnt, uint32_t* __restrict__ vmask,
                uint32_t* __restrict__ tile_sites,
                uint32_t* __restrict__ tile_faces)
{
    __shared__ uint32_t sh[kBlock / 64];
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t ns = 0, nf = 0;
    Chunk K;
    if (load_chunk(occ, D, c, K))
    {
        uint32_t em[kVz];
        const uint32_t cnt = decide_chunk(K, D, em);
        ns = cnt >> 16;
        nf = cnt & 0xFFFFu;
        chunk_cnt[c] = cnt;
        if (cnt)
        {
            uint4* dst = reinterpret_cast<uint4*>(vmask + (int64_t)c * kVz);
            dst[0] = make_uint4(em[0], em[1], em[2], em[3]);
            dst[1] = make_uint4(em[4], em[5], em[6], em[7]);
        }
    }
    uint32_t tot;
    block_scan_excl((ns << 16) | nf, sh, &tot);
    if (threadIdx.x == 0)
    {
        tile_sites[blockIdx.x] = tot >> 16;
        tile_faces[blockIdx.x] = tot & 0xFFFFu;
    }
}

// chunk c of the vectorised passes: row (x, y) and first z
__device__ __forceinline__ void chunk_at(const Dims& D, int64_t c, int& x, int& y,
                                         int& z0)
{
    const int nzc = D.nz / kVz;
    const int64_t row = c / nzc;
    z0 = (int)(c - row * nzc) * kVz;
    x = (int)(row / D.ny);
    y = (int)(row - (int64_t)x * D.ny);
}

__global__ void __launch_bounds__(kBlock)
    bnd_emit_sites_v(Dims D, const uint32_t* __restrict__ chunk_cnt,
                     const uint32_t* __restrict__ vmask,
                     const uint32_t* __restrict__ tile_sites,
                     float* __restrict__ sites, uint32_t* __restrict__ cmap)
{
    __shared__ uint32_t sh[kBlock / 64];
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nch = (int64_t)D.nx * D.ny * (D.nz / kVz);
    const uint32_t ns = c < nch ? chunk_cnt[c] >> 16 : 0u;
    uint32_t tot;
    uint32_t s = tile_sites[blockIdx.x] + block_scan_excl(ns, sh, &tot);
    if (ns)
    {
        int x, y, z0;
        chunk_at(D, c, x, y, z0);
        const uint4* src = reinterpret_cast<const uint4*>(vmask + c * kVz);
        const uint4 a = src[0], b = src[1];
        const uint32_t em[kVz] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        for (int j = 0; j < kVz; ++j)
        {
            const Emit e{em[j] & 0xFFFFFFu, 0u};
            put_sites(e, D, x, y, z0 + j, s, sites, cmap);
            s += (uint32_t)__builtin_popcount(e.keys);
        }
    }
}

__global__ void __launch_bounds__(kBlock)
    bnd_emit_quads_v(Dims D, const uint32_t* __restrict__ chunk_cnt,
                     const uint32_t* __restrict__ vmask,
                     const uint32_t* __restrict__ tile_faces,
                     const uint32_t* __restrict__ cmap,
                     int32_t* __restrict__ quads)
{
    __shared__ uint32_t sh[kBlock / 64];
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nch = (int64_t)D.nx * D.ny * (D.nz / kVz);
    const uint32_t nf = c < nch ? chunk_cnt[c] & 0xFFFFu : 0u;
    uint32_t tot;
    uint32_t f = tile_faces[blockIdx.x] + block_scan_excl(nf, sh, &tot);
    if (nf)
    {
        int x, y, z0;
        chunk_at(D, c, x, y, z0);
        const uint4* src = reinterpret_cast<const uint4*>(vmask + c * kVz);
        const uint4 a = src[0], b = src[1];
        const uint32_t em[kVz] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        for (int j = 0; j < kVz; ++j)
        {
            const Emit e{0u, em[j] >> 24};
            put_quads(e, D, x, y, z0 + j, f, cmap, quads);
            f += (uint32_t)__builtin_popcount(e.faces);
        }
    }
}

// exclusive scan of a uint32 array in place, one block of 1024 per chunk;
// writes the chunk totals to `sums`
__global__ void __launch_bounds__(1024)
    scan_chunks(uint32_t* a, int64_t n, uint32_t* sums)
{
    __shared__ uint32_t sh[16];
    const int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    const uint32_t v = i < n ? a[i] : 0u;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1)
    {
        const uint32_t t = __shfl_up(incl, d);
        if (lane >= d)
            incl += t;
    }
    if (lane == 63)
        sh[wv] = incl;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int w = 0; w < 16; ++w)
    {
        if (w < wv)
            off += sh[w];
        tot += sh[w];
    }
    if (i < n)
        a[i] = off + incl - v;
    if (threadIdx.x == 0)
        sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(1024)
    scan_add(uint32_t* a, int64_t n, const uint32_t* sums)
{
    const int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    if (i < n)
        a[i] += sums[blockIdx.x];
}

// pass 3: sites (coordinates + corner map)
__global__ void __launch_bounds__(kBlock)
    bnd_emit_sites(const uint8_t* __restrict__ occ, Dims D, int64_t N,
                   const uint32_t* __restrict__ tile_sites,
                   float* __restrict__ sites, uint32_t* __restrict__ cmap)
{
    __shared__ uint32_t sh[kBlock / 64];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    Emit e{0u, 0u};
    int x = 0, y = 0, z = 0;
    if (i < N)
    {
        decode(D, i, x, y, z);
        if (face_event(occ, D, x, y, z))
            e = decide(gather(occ, D, x, y, z));
    }
    uint32_t tot;
    const uint32_t base = tile_sites[blockIdx.x] +
                          block_scan_excl((uint32_t)__builtin_popcount(e.keys), sh, &tot);
    put_sites(e, D, x, y, z, base, sites, cmap);
}

// pass 4: quads through the corner map
__global__ void __launch_bounds__(kBlock)
    bnd_emit_quads(const uint8_t* __restrict__ occ, Dims D, int64_t N,
                   const uint32_t* __restrict__ tile_faces,
                   const uint32_t* __restrict__ cmap, int32_t* __restrict__ quads)
{
    __shared__ uint32_t sh[kBlock / 64];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    Emit e{0u, 0u};
    int x = 0, y = 0, z = 0;
    if (i < N)
    {
        decode(D, i, x, y, z);
        if (face_event(occ, D, x, y, z))
            e = decide(gather(occ, D, x, y, z));
    }
    uint32_t tot;
    const uint32_t base = tile_faces[blockIdx.x] +
                          block_scan_excl((uint32_t)__builtin_popcount(e.faces), sh, &tot);
    put_quads(e, D, x, y, z, base, cmap, quads);
}

// ------------------------------------------------------------- Euler terms
// accumulators: [0] V_bnd, [1] E_bnd, [2] Vint_sum, [3] E_int, [4] F_int,
// [5] T, [6] thin-layer components, [7] union-find guard, [8] / [9] positive
// / negative terms of the cubical Euler characteristic, spread over 64 slots
constexpr int kSlots = 64;
constexpr int kAcc = 11;

__device__ __forceinline__ void block_add(unsigned long long v,
                                          unsigned long long* sh,
                                          unsigned long long* dst)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        v += __shfl_xor(v, d);
    const int wv = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
        sh[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned long long t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
            t += sh[w];
        if (t)
            atomicAdd(dst, t);
    }
}

__device__ __forceinline__ int occ_at(const uint8_t* occ, const Dims& D, int x,
                                      int y, int z)
{
    if ((unsigned)x >= (unsigned)D.nx || (unsigned)y >= (unsigned)D.ny ||
        (unsigned)z >= (unsigned)D.nz)
        return 0;
    return occ[((int64_t)x * D.ny + y) * D.nz + z] != 0;
}

// Euler characteristic of the union of the closed voxel cubes (cubical
// complex, the topological invariant of the 26-connected object): per
// lattice point p, with m the inside bits of the 8 voxels p + (a, b, c),
// a, b, c in {-1, 0} (bit (a+1)*4 + (b+1)*2 + (c+1)): the vertex p, the
// edges and faces leaving p towards +x, +y, +z and the voxel at p each exist
// iff one of their incident voxels is inside.
__global__ void __launch_bounds__(kBlock)
    chi_lattice(const uint8_t* __restrict__ occ, Dims D,
                unsigned long long* acc)
{
    __shared__ unsigned long long sh[kBlock / 64];
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, x = blockIdx.z;
    unsigned long long pos = 0, neg = 0;
    if (z <= D.nz)
    {
        uint32_t m = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            m |= (uint32_t)occ_at(occ, D, x + (q >> 2) - 1,
                                  y + ((q >> 1) & 1) - 1, z + (q & 1) - 1)
                 << q;
        pos = (m != 0) + ((m & 0xC0u) != 0) + ((m & 0xA0u) != 0) +
              ((m & 0x88u) != 0);
        neg = ((m & 0xF0u) != 0) + ((m & 0xCCu) != 0) + ((m & 0xAAu) != 0) +
              ((m & 0x80u) != 0);
    }
    const int slot = (int)((blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y *
                                                       (uint64_t)blockIdx.z)) %
                           kSlots);
    block_add(pos, sh, &acc[slot * kAcc + 8]);
    block_add(neg, sh, &acc[slot * kAcc + 9]);
}

// per corner-lattice point (cx, cy, cz): the boundary-vertex multiplicity if
// it is a site, and the boundary / interior edges starting at it (+x, +y, +z)
__global__ void __launch_bounds__(kBlock)
    euler_lattice(const uint8_t* __restrict__ occ, Dims D,
                  const int32_t* __restrict__ tbl, unsigned long long* acc)
{
    __shared__ unsigned long long sh[kBlock / 64];
    const int cz = blockIdx.x * blockDim.x + threadIdx.x;
    const int cy = blockIdx.y, cx = blockIdx.z;
    unsigned long long vb = 0, eb = 0, ei = 0;
    if (cz <= D.nz)
    {
        int blk = 0; // occupancy of the 8 voxels around the corner
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int c = 0; c < 2; ++c)
                    blk += occ_at(occ, D, cx - 1 + a, cy - 1 + b, cz - 1 + c);
        if (blk == 0 || blk == 8)
        {
            // uniform block: no site, no boundary edge; every edge from this
            // point inside the grid is interior iff the block is full
            if (blk == 8)
                ei = (cx < D.nx) + (cy < D.ny) + (cz < D.nz);
        }
        else
        {
            // site: mask around the (int)-truncated base voxel
            // (include/spaceinfo.h:140-153), bit (7 - i) for offset i
            const int bx = cx > 0 ? cx - 1 : 0, by = cy > 0 ? cy - 1 : 0,
                      bz = cz > 0 ? cz - 1 : 0;
            int mask = 0;
#pragma unroll
            for (int i = 0; i < 8; ++i)
                mask |= occ_at(occ, D, bx + ((i >> 2) & 1), by + ((i >> 1) & 1),
                               bz + (i & 1))
                        << (7 - i);
            vb = (unsigned long long)tbl[mask];
        // edges (+x, +y, +z) from this lattice point: the 4 voxels around
        // the edge, in cyclic order; boundary faces = value changes
#pragma unroll
        for (int ax = 0; ax < 3; ++ax)
        {
            const int lim = ax == 0 ? D.nx : (ax == 1 ? D.ny : D.nz);
            const int cc = ax == 0 ? cx : (ax == 1 ? cy : cz);
            if (cc >= lim)
                continue;
            int v4[4];
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                const int u = (q == 1 || q == 2) ? 0 : -1; // cyclic square
                const int w = (q >= 2) ? 0 : -1;
                int p[3] = {cx, cy, cz};
                p[(ax + 1) % 3] += u;
                p[(ax + 2) % 3] += w;
                v4[q] = occ_at(occ, D, p[0], p[1], p[2]);
            }
            const int ch = (v4[0] != v4[1]) + (v4[1] != v4[2]) +
                           (v4[2] != v4[3]) + (v4[3] != v4[0]);
            if (ch == 2)
                eb += 1;
            else if (ch == 4)
                eb += 2;
            if (v4[0] + v4[1] + v4[2] + v4[3] == 4)
                ei += 1;
        }
        }
    }
    const int slot = (blockIdx.x + blockIdx.y * 3 + blockIdx.z * 7) % kSlots;
    unsigned long long* a = acc + slot * kAcc;
    block_add(vb, sh, &a[0]);
    block_add(eb, sh, &a[1]);
    block_add(ei, sh, &a[3]);
}

// per voxel: T, F_int and the reference's interior-corner sum
__global__ void __launch_bounds__(kBlock)
    euler_voxels(const uint8_t* __restrict__ occ, Dims D,
                 unsigned long long* acc)
{
    __shared__ unsigned long long sh[kBlock / 64];
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, x = blockIdx.z;
    unsigned long long t = 0, fi = 0, vi = 0;
    if (z < D.nz && occ_at(occ, D, x, y, z))
    {
        t = 1;
        // one gather of the 3 x 3 x 3 neighbourhood (outside the grid = 0)
        const uint32_t m = gather(occ, D, x, y, z).occ;
        fi = obit(m, -1, 0, 0) + obit(m, 1, 0, 0) + obit(m, 0, -1, 0) +
             obit(m, 0, 1, 0) + obit(m, 0, 0, -1) + obit(m, 0, 0, 1);
        // vox_corner_offset order; base voxel = (int)(coord +- 0.5): the
        // block starts at offset -1, or 0 on the high side or at coordinate 0
#pragma unroll
        for (int c = 0; c < 8; ++c)
        {
            const int ax = (((c >> 2) & 1) || x == 0) ? 0 : -1;
            const int ay = (((c >> 1) & 1) || y == 0) ? 0 : -1;
            const int az = ((c & 1) || z == 0) ? 0 : -1;
            uint32_t bm = 0;
#pragma unroll
            for (int i = 0; i < 8; ++i)
                bm |= 1u << nbit(ax + ((i >> 2) & 1), ay + ((i >> 1) & 1),
                                 az + (i & 1));
            vi += (m & bm) == bm;
        }
    }
    const int slot = (blockIdx.x + blockIdx.y * 3 + blockIdx.z * 7) % kSlots;
    unsigned long long* a = acc + slot * kAcc;
    block_add(vi, sh, &a[2]);
    block_add(fi, sh, &a[4]);
    block_add(t, sh, &a[5]);
}

// --------------------------------------------- thin layer + union-find (C)
__global__ void __launch_bounds__(kBlock)
    layer_boundary(const uint8_t* __restrict__ occ, Dims D,
                   uint8_t* __restrict__ bnd)
{
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, x = blockIdx.z;
    if (z >= D.nz)
        return;
    uint8_t b = 0;
    if (occ_at(occ, D, x, y, z))
        b = !(occ_at(occ, D, x - 1, y, z) && occ_at(occ, D, x + 1, y, z) &&
              occ_at(occ, D, x, y - 1, z) && occ_at(occ, D, x, y + 1, z) &&
              occ_at(occ, D, x, y, z - 1) && occ_at(occ, D, x, y, z + 1));
    bnd[((int64_t)x * D.ny + y) * D.nz + z] = b;
}

__global__ void __launch_bounds__(kBlock)
    layer_init(const uint8_t* __restrict__ occ, const uint8_t* __restrict__ bnd,
               Dims D, int32_t* __restrict__ lab)
{
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, x = blockIdx.z;
    if (z >= D.nz)
        return;
    const int64_t i = ((int64_t)x * D.ny + y) * D.nz + z;
    bool in = false;
    if (occ[i])
        for (int dx = -1; dx <= 1 && !in; ++dx)
            for (int dy = -1; dy <= 1 && !in; ++dy)
                for (int dz = -1; dz <= 1 && !in; ++dz)
                {
                    const int a = x + dx, b = y + dy, c = z + dz;
                    if ((unsigned)a < (unsigned)D.nx &&
                        (unsigned)b < (unsigned)D.ny &&
                        (unsigned)c < (unsigned)D.nz &&
                        bnd[((int64_t)a * D.ny + b) * D.nz + c])
                        in = true;
                }
    lab[i] = in ? (int32_t)i : -1;
}

__device__ __forceinline__ int32_t uf_find(int32_t* lab, int32_t i)
{
    int32_t p = lab[i];
    while (p != i)
    {
        const int32_t g = lab[p];
        if (g == p)
            break;
        // path halving (benign races: parents only ever decrease)
        lab[i] = g;
        i = p;
        p = g;
    }
    return p;
}

__device__ void uf_union(int32_t* lab, int32_t a, int32_t b,
                         unsigned long long* guard_hits)
{
    // the smaller root wins; a CAS loop makes the link atomic (lock-free: a
    // failed CAS means another link succeeded); the bound is a safety net
    for (int guard = 0; guard < (1 << 20); ++guard)
    {
        a = uf_find(lab, a);
        b = uf_find(lab, b);
        if (a == b)
            return;
        if (a > b)
        {
            const int32_t t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(&lab[b], b, a) == b)
            return;
    }
    atomicAdd(guard_hits, 1ull);
}

__global__ void __launch_bounds__(kBlock)
    layer_union(int32_t* lab, Dims D, unsigned long long* guard_hits)
{
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, x = blockIdx.z;
    if (z >= D.nz)
        return;
    const int64_t sy = D.nz, sx = (int64_t)D.ny * D.nz;
    const int64_t i = x * sx + y * sy + z;
    if (lab[i] < 0)
        return;
    if (x + 1 < D.nx && lab[i + sx] >= 0)
        uf_union(lab, (int32_t)i, (int32_t)(i + sx), guard_hits);
    if (y + 1 < D.ny && lab[i + sy] >= 0)
        uf_union(lab, (int32_t)i, (int32_t)(i + sy), guard_hits);
    if (z + 1 < D.nz && lab[i + 1] >= 0)
        uf_union(lab, (int32_t)i, (int32_t)(i + 1), guard_hits);
}

__global__ void __launch_bounds__(kBlock)
    layer_count(const int32_t* __restrict__ lab, int64_t N,
                unsigned long long* acc)
{
    __shared__ unsigned long long sh[kBlock / 64];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned long long r = (i < N && lab[i] == (int32_t)i) ? 1ull : 0ull;
    block_add(r, sh, &acc[(blockIdx.x % kSlots) * kAcc + 6]);
}

// ---- vectorised Euler terms and thin layer (nz a multiple of 8).  The same
// 10-bit row windows as the site / quad passes; one thread per 8 consecutive
// z of a row, so every neighbour read of the per-voxel kernels above becomes
// a bit of one of nine windows (27 loads per 8 voxels instead of up to 27
// per voxel).

// block sum of the eight Euler terms, one atomic per term and block
__device__ __forceinline__ void block_add8(const uint32_t (&v)[8],
                                           unsigned long long (*sh)[kBlock / 64],
                                           unsigned long long* a)
{
    const int wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 8; ++k)
    {
        uint32_t s = v[k];
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1)
            s += __shfl_xor(s, d);
        if ((threadIdx.x & 63) == 0)
            sh[k][wv] = s;
    }
    __syncthreads();
    if (threadIdx.x < 8)
    {
        unsigned long long t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
            t += sh[threadIdx.x][w];
        // accumulator of term k: V_bnd, E_bnd, Vint_sum, E_int, F_int, T,
        // chi positive, chi negative
        const int dst = threadIdx.x < 6 ? (int)threadIdx.x : (int)threadIdx.x + 2;
        if (t)
            atomicAdd(&a[dst], t);
    }
}

// Lattice row (cx, cy) in [0, nx] x [0, ny], lattice points z0 .. z0 + 7 (and
// nz on the row's last chunk); voxel (cx, cy, z) for the same z when the row
// is a voxel row.  Window w[dx][dy] is voxel row (cx + dx - 1, cy + dy - 1)
// and its bit b is voxel z0 - 1 + b, so lattice point cz = z0 + j sees
// voxels cz - 1, cz at bits j, j + 1.  Terms and conventions as chi_lattice,
// euler_lattice and euler_voxels, evaluated bit-sliced: each quantity is a
// word whose bit j belongs to lattice point (or voxel) z0 + j, so all points
// of the chunk are done by the same few bitwise operations; only the
// boundary-corner table lookups run per point.
__global__ void __launch_bounds__(kBlock)
    euler_rows_v(const uint8_t* __restrict__ occ, Dims D,
                 const int32_t* __restrict__ tbl, unsigned long long* acc)
{
    __shared__ unsigned long long sh[8][kBlock / 64];
    __shared__ int32_t stbl[256];
    for (int i = threadIdx.x; i < 256; i += blockDim.x)
        stbl[i] = tbl[i];
    __syncthreads();
    const int nzc = D.nz / kVz;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // V_bnd, E_bnd, Vint_sum, E_int, F_int, T, chi+, chi-
    uint32_t s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // lattice rows in linear order, chunks of a row consecutive: the windows
    // are read by every lane (row_win shuffles), the terms by live ones
    const bool live = c < (int64_t)(D.nx + 1) * (D.ny + 1) * nzc;
    const int64_t row = (live ? c : 0) / nzc;
    const int z0 = (int)((live ? c : 0) - row * nzc) * kVz;
    const int cx = (int)(row / (D.ny + 1));
    const int cy = (int)(row - (int64_t)cx * (D.ny + 1));
    uint32_t w[3][3];
#pragma unroll
    for (int dx = 0; dx < 3; ++dx)
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
            w[dx][dy] = row_win(occ, D, cx + dx - 1, cy + dy - 1, z0, live);
    if (live)
    {
        // lattice points of the chunk (a ninth, cz = nz, on a row's last)
        const uint32_t JM = z0 + kVz == D.nz ? 0x1FFu : 0xFFu;
        const uint32_t zf = z0 == 0 ? 1u : 0u; // point / voxel at z = 0
        // voxel (cx - 1 + a, cy - 1 + b, cz - 1 + e) of lattice point cz
        auto L = [&](int a, int b, int e) -> uint32_t { return (w[a][b] >> e) & JM; };
        const uint32_t l000 = L(0, 0, 0), l001 = L(0, 0, 1), l010 = L(0, 1, 0),
                       l011 = L(0, 1, 1), l100 = L(1, 0, 0), l101 = L(1, 0, 1),
                       l110 = L(1, 1, 0), l111 = L(1, 1, 1);
        const uint32_t any = l000 | l001 | l010 | l011 | l100 | l101 | l110 | l111;
        const uint32_t all = l000 & l001 & l010 & l011 & l100 & l101 & l110 & l111;
        s[6] = __builtin_popcount(any) + __builtin_popcount(l110 | l111) +
               __builtin_popcount(l101 | l111) + __builtin_popcount(l011 | l111);
        s[7] = __builtin_popcount(l100 | l101 | l110 | l111) +
               __builtin_popcount(l010 | l011 | l110 | l111) +
               __builtin_popcount(l001 | l011 | l101 | l111) + __builtin_popcount(l111);
        // edges from the point along +x, +y, +z (inside the grid): the four
        // voxels around each, in cyclic order
        auto edge = [&](uint32_t v0, uint32_t v1, uint32_t v2, uint32_t v3, uint32_t lim) {
            const uint32_t d0 = v0 ^ v1, d1 = v1 ^ v2, d2 = v2 ^ v3, d3 = v3 ^ v0;
            const uint32_t ch4 = d0 & d1 & d2 & d3;
            const uint32_t ch2 = (d0 | d1 | d2 | d3) & ~ch4;
            s[1] += __builtin_popcount(ch2 & lim) + 2 * __builtin_popcount(ch4 & lim);
            s[3] += __builtin_popcount(v0 & v1 & v2 & v3 & lim);
        };
        edge(l100, l110, l111, l101, cx < D.nx ? JM : 0u);
        edge(l010, l011, l111, l110, cy < D.ny ? JM : 0u);
        edge(l001, l101, l111, l011, 0xFFu);
        // boundary corners (some but not all of the 8 voxels inside): the
        // table of the site mask, whose block starts at voxel 0 where the
        // coordinate is 0
        const uint32_t part = any & ~all;
        if (part)
        {
            const bool bx = cx == 0, by = cy == 0;
            uint64_t X = 0;
            uint32_t last = 0;
#pragma unroll
            for (int i = 0; i < 8; ++i)
            {
                const int ix = (i >> 2) & 1, iy = (i >> 1) & 1, iz = i & 1;
                // selects, not a run-time index (which would put w in memory)
                const uint32_t r = bx ? (by ? w[ix + 1][iy + 1] : w[ix + 1][iy])
                                      : (by ? w[ix][iy + 1] : w[ix][iy]);
                const uint32_t si = (((r >> iz) & ~zf) | ((r >> (iz + 1)) & zf)) & JM;
                X |= (uint64_t)(si & 0xFFu) << (8 * (7 - i));
                last |= ((si >> 8) & 1u) << (7 - i);
            }
            const uint64_t T = transpose8(X);
            for (uint32_t p = part; p; p &= p - 1)
            {
                const int j = __builtin_ctz(p);
                const uint32_t mask = j < 8 ? (uint32_t)(T >> (8 * j)) & 0xFFu : last;
                s[0] += (uint32_t)stbl[mask];
            }
        }
        // inside voxels (cx, cy, z0 + j): T, F_int and the interior-corner sum
        if (cx < D.nx && cy < D.ny)
        {
            auto V = [&](int dx, int dy, int dz) -> uint32_t {
                return (w[dx + 1][dy + 1] >> (dz + 1)) & 0xFFu;
            };
            const uint32_t in = V(0, 0, 0);
            s[5] = __builtin_popcount(in);
            s[4] = __builtin_popcount(in & V(-1, 0, 0)) + __builtin_popcount(in & V(1, 0, 0)) +
                   __builtin_popcount(in & V(0, -1, 0)) + __builtin_popcount(in & V(0, 1, 0)) +
                   __builtin_popcount(in & V(0, 0, -1)) + __builtin_popcount(in & V(0, 0, 1));
            // 2x2x2 block (ax..ax+1, ay..ay+1, az..az+1) all inside
            auto B = [&](int ax, int ay, int az) -> uint32_t {
                return V(ax, ay, az) & V(ax, ay, az + 1) & V(ax, ay + 1, az) &
                       V(ax, ay + 1, az + 1) & V(ax + 1, ay, az) & V(ax + 1, ay, az + 1) &
                       V(ax + 1, ay + 1, az) & V(ax + 1, ay + 1, az + 1);
            };
            const uint32_t b000 = B(-1, -1, -1), b001 = B(-1, -1, 0), b010 = B(-1, 0, -1),
                           b011 = B(-1, 0, 0), b100 = B(0, -1, -1), b101 = B(0, -1, 0),
                           b110 = B(0, 0, -1), b111 = B(0, 0, 0);
            // block with offset 0 (true) or -1 (false) per axis: selects of
            // values, so the windows stay in registers
            auto blk = [&](bool X0, bool Y0, bool Z0) -> uint32_t {
                return X0 ? (Y0 ? (Z0 ? b111 : b110) : (Z0 ? b101 : b100))
                          : (Y0 ? (Z0 ? b011 : b010) : (Z0 ? b001 : b000));
            };
            // corner slot q: block offset -1, or 0 on its high side or at
            // coordinate 0 (in z only the voxel at z = 0)
#pragma unroll
            for (int q = 0; q < 8; ++q)
            {
                const bool X0 = ((q >> 2) & 1) || cx == 0;
                const bool Y0 = ((q >> 1) & 1) || cy == 0;
                const uint32_t full = (q & 1) ? blk(X0, Y0, true)
                                              : ((blk(X0, Y0, false) & ~zf) |
                                                 (blk(X0, Y0, true) & zf));
                s[2] += __builtin_popcount(in & full);
            }
        }
    }
    block_add8(s, sh, acc + (blockIdx.x % kSlots) * kAcc);
}

// boundary voxels of 8 consecutive z: inside with a non-inside face
// neighbour (outside the grid counts as non-inside), one byte per chunk
// (bit j = voxel z0 + j)
__global__ void __launch_bounds__(kBlock)
    layer_boundary_v(const uint8_t* __restrict__ occ, Dims D,
                     uint8_t* __restrict__ bnd8)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    Chunk K;
    if (!load_chunk(occ, D, c, K))
        return;
    const uint32_t me = K.w[1][1] >> 1;
    const uint32_t all6 = (K.w[1][1] & (K.w[1][1] >> 2)) & (K.w[0][1] >> 1) &
                          (K.w[2][1] >> 1) & (K.w[1][0] >> 1) & (K.w[1][2] >> 1);
    bnd8[c] = (uint8_t)(me & ~all6 & 0xFFu);
}

// ---- thin layer, vectorised path: two-level labelling over voxel positions.
// lay8[c] marks the layer voxels of chunk c (inside and within one voxel,
// 26-wise, of a boundary voxel).  A tile of kTX x kTY x kTZ chunks (8 x 8 x 32
// voxels, one workgroup of 256) is labelled in LDS: local ids are positions in the tile (chunk
// t, bit j -> 8 t + j), runs within a chunk start as trees, and the pairs
// inside the tile are joined by a lock-free union-find on LDS; every layer
// voxel's label (its linear voxel index in lab) is then the linear index of
// its tile-local root.  A second pass joins the pairs that cross tile faces
// on those labels.  Components = tile-local roots - successful cross-tile
// links (each link joins two components), so no pass counts roots.
constexpr int kTX = 8, kTY = 8, kTZ = 4, kTile = kTX * kTY * kTZ;

// 10-bit window of the boundary marks of row (x, y) at chunk z0 (bit b =
// voxel z0 - 1 + b, outside the grid = 0)
__device__ __forceinline__ uint32_t bnd_win(const uint8_t* __restrict__ bnd8,
                                            const Dims& D, int x, int y, int z0)
{
    if ((unsigned)x >= (unsigned)D.nx || (unsigned)y >= (unsigned)D.ny)
        return 0;
    const int64_t c = ((int64_t)x * D.ny + y) * (D.nz / kVz) + z0 / kVz;
    const uint32_t lo = z0 > 0 ? (uint32_t)bnd8[c - 1] >> 7 : 0u;
    const uint32_t hi = z0 + kVz < D.nz ? (uint32_t)bnd8[c + 1] & 1u : 0u;
    return lo | ((uint32_t)bnd8[c] << 1) | (hi << 9);
}

__device__ __forceinline__ int32_t lds_find(int32_t* par, int32_t i)
{
    int32_t p = par[i];
    while (p != i)
    {
        const int32_t g = par[p];
        if (g == p)
            break;
        par[i] = g;
        i = p;
        p = g;
    }
    return p;
}

__device__ __forceinline__ void lds_union(int32_t* par, int32_t a, int32_t b)
{
    // terminates: every failed CAS is another thread's successful link, and
    // there are fewer links than tile voxels
    for (;;)
    {
        a = lds_find(par, a);
        b = lds_find(par, b);
        if (a == b)
            return;
        if (a > b)
        {
            const int32_t t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(&par[b], b, a) == b)
            return;
    }
}

// the x / y pairs of chunks `me` and `nb` (same z) to join: a pair whose
// predecessor in z is also a pair is already joined through it (runs within
// a chunk are trees)
__device__ __forceinline__ uint32_t pair_starts(uint32_t me, uint32_t nb)
{
    const uint32_t m = me & nb;
    return m & ~(m << 1);
}

__global__ void __launch_bounds__(kTile)
    layer_tile(const uint8_t* __restrict__ occ, const uint8_t* __restrict__ bnd8,
               Dims D, int tiles_y, int tiles_z, int64_t n_tiles,
               uint8_t* __restrict__ lay8, int32_t* __restrict__ lab,
               unsigned long long* acc)
{
    __shared__ int32_t par[kTile * kVz];
    __shared__ uint8_t m8[kTile];
    __shared__ unsigned long long sh[kTile / 64];
    // XCD-mapped: tiles next to each other in z (which read the same lines)
    // go to one XCD's L2
    const uint32_t item = vc_xcd_item(blockIdx.x, gridDim.x / 8);
    if (item >= (uint32_t)n_tiles)
        return;
    const int t = threadIdx.x;
    const int tz = t % kTZ, ty = (t / kTZ) % kTY, tx = t / (kTZ * kTY);
    int blk = (int)item;
    const int bz = blk % tiles_z;
    blk /= tiles_z;
    const int by = blk % tiles_y, bx = blk / tiles_y;
    const int nzc = D.nz / kVz;
    const int x = bx * kTX + tx, y = by * kTY + ty, zc = bz * kTZ + tz;
    const bool in = x < D.nx && y < D.ny && zc < nzc;
    uint32_t me = 0;
    if (in)
    {
        const int z0 = zc * kVz;
        uint32_t near = 0; // bit b: a boundary voxel within z0 - 1 + b +- 1
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy)
            {
                const uint32_t w = bnd_win(bnd8, D, x + dx, y + dy, z0);
                near |= w | (w >> 1) | (w << 1);
            }
        const uint64_t o8 = *reinterpret_cast<const uint64_t*>(
            occ + ((int64_t)x * D.ny + y) * D.nz + z0);
#pragma unroll
        for (int j = 0; j < kVz; ++j)
            me |= (((o8 >> (8 * j)) & 0xFFu) != 0 && ((near >> (j + 1)) & 1u)) ? 1u << j : 0u;
        lay8[((int64_t)x * D.ny + y) * nzc + zc] = (uint8_t)me;
    }
    m8[t] = (uint8_t)me;
    if (me)
    {
        // runs within the chunk rooted at their first voxel (entries of
        // voxels outside the layer are never read)
        int32_t start = t * kVz;
#pragma unroll
        for (int j = 0; j < kVz; ++j)
        {
            if ((me >> j) & 1u)
                par[t * kVz + j] = start;
            else
                start = t * kVz + j + 1;
        }
    }
    __syncthreads();
    if (me)
    {
        if (tz + 1 < kTZ && ((me >> 7) & 1u) && (m8[t + 1] & 1u))
            lds_union(par, t * kVz + kVz - 1, (t + 1) * kVz);
        if (tx + 1 < kTX)
            for (uint32_t u = pair_starts(me, m8[t + kTZ * kTY]); u; u &= u - 1)
            {
                const int j = __builtin_ctz(u);
                lds_union(par, t * kVz + j, (t + kTZ * kTY) * kVz + j);
            }
        if (ty + 1 < kTY)
            for (uint32_t u = pair_starts(me, m8[t + kTZ]); u; u &= u - 1)
            {
                const int j = __builtin_ctz(u);
                lds_union(par, t * kVz + j, (t + kTZ) * kVz + j);
            }
    }
    __syncthreads();
    unsigned long long roots = 0;
    if (me)
    {
        const int64_t base = ((int64_t)x * D.ny + y) * D.nz + zc * kVz;
        for (uint32_t u = me; u; u &= u - 1)
        {
            const int j = __builtin_ctz(u);
            const int32_t r = lds_find(par, t * kVz + j);
            roots += r == t * kVz + j ? 1ull : 0ull;
            const int rt = r / kVz, rj = r % kVz;
            const int rx = bx * kTX + rt / (kTZ * kTY), ry = by * kTY + (rt / kTZ) % kTY;
            const int rz = (bz * kTZ + rt % kTZ) * kVz + rj;
            lab[base + j] = (int32_t)(((int64_t)rx * D.ny + ry) * D.nz + rz);
        }
    }
    block_add(roots, sh, &acc[(blockIdx.x % kSlots) * kAcc + 6]);
}

// union that reports whether it joined two components
__device__ uint32_t uf_link(int32_t* lab, int32_t a, int32_t b,
                            unsigned long long* guard_hits)
{
    for (int guard = 0; guard < (1 << 20); ++guard)
    {
        a = uf_find(lab, a);
        b = uf_find(lab, b);
        if (a == b)
            return 0;
        if (a > b)
        {
            const int32_t t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(&lab[b], b, a) == b)
            return 1;
    }
    atomicAdd(guard_hits, 1ull);
    return 0;
}

// the pairs that cross a tile face (+x, +y, +z); successful links counted
__global__ void __launch_bounds__(kBlock)
    layer_border(const uint8_t* __restrict__ lay8, Dims D, int32_t* lab,
                 unsigned long long* acc)
{
    __shared__ unsigned long long sh[kBlock / 64];
    const int nzc = D.nz / kVz;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t me = c < (int64_t)D.nx * D.ny * nzc ? lay8[c] : 0u;
    unsigned long long links = 0;
    if (me)
    {
        unsigned long long* guard = acc + 7;
        const int64_t row = c / nzc;
        const int zc = (int)(c - row * nzc);
        const int x = (int)(row / D.ny);
        const int y = (int)(row - (int64_t)x * D.ny);
        const int32_t p = (int32_t)(row * D.nz + zc * kVz);
        if ((x + 1) % kTX == 0 && x + 1 < D.nx)
        {
            const int32_t sx = D.ny * D.nz;
            for (uint32_t u = pair_starts(me, lay8[c + (int64_t)D.ny * nzc]); u; u &= u - 1)
            {
                const int j = __builtin_ctz(u);
                links += uf_link(lab, p + j, p + j + sx, guard);
            }
        }
        if ((y + 1) % kTY == 0 && y + 1 < D.ny)
            for (uint32_t u = pair_starts(me, lay8[c + nzc]); u; u &= u - 1)
            {
                const int j = __builtin_ctz(u);
                links += uf_link(lab, p + j, p + j + D.nz, guard);
            }
        if ((zc + 1) % kTZ == 0 && zc + 1 < nzc && ((me >> 7) & 1u) && (lay8[c + 1] & 1u))
            links += uf_link(lab, p + kVz - 1, p + kVz, guard);
    }
    block_add(links, sh, &acc[(blockIdx.x % kSlots) * kAcc + 10]);
}

// the reference's per-mask surface-component table, derived from first
// principles (6-components of the inside corners + 18-components of the
// outside corners - 1; four antipodal hole masks and 0/255 give 0)
void cube_table(int32_t* tbl)
{
    auto comps = [](int set, bool adj18) {
        int seen = 0, n = 0;
        for (int s = 0; s < 8; ++s)
        {
            if (!((set >> s) & 1) || ((seen >> s) & 1))
                continue;
            ++n;
            int stack[8], sp = 0;
            stack[sp++] = s;
            seen |= 1 << s;
            while (sp)
            {
                const int u = stack[--sp];
                for (int v = 0; v < 8; ++v)
                {
                    if (!((set >> v) & 1) || ((seen >> v) & 1))
                        continue;
                    const int pc = __builtin_popcount(u ^ v);
                    if (pc == 1 || (adj18 && pc == 2))
                    {
                        seen |= 1 << v;
                        stack[sp++] = v;
                    }
                }
            }
        }
        return n;
    };
    for (int mask = 0; mask < 256; ++mask)
    {
        int in = 0, out = 0;
        for (int i = 0; i < 8; ++i)
            if ((mask >> (7 - i)) & 1)
                in |= 1 << i;
            else
                out |= 1 << i;
        tbl[mask] = (mask == 0 || mask == 255)
                        ? 0
                        : comps(in, false) + comps(out, true) - 1;
    }
    tbl[126] = tbl[189] = tbl[219] = tbl[231] = 0;
}

int scan_inplace(vc_ctx* ctx, uint32_t* a, int64_t n)
{
    // exclusive scan with chunks of 1024, recursing on the chunk totals
    const int64_t chunks = (n + 1023) / 1024;
    uint32_t* sums = nullptr;
    int rc = vc_alloc(ctx, (void**)&sums, 4 * (size_t)(chunks > 0 ? chunks : 1));
    if (rc != VC_OK)
        return rc;
    scan_chunks<<<(unsigned)chunks, 1024, 0, ctx->stream>>>(a, n, sums);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && chunks > 1)
    {
        rc = scan_inplace(ctx, sums, chunks);
        if (rc == VC_OK)
        {
            scan_add<<<(unsigned)chunks, 1024, 0, ctx->stream>>>(a, n, sums);
            e = hipGetLastError();
        }
    }
    hipError_t e2 = hipStreamSynchronize(ctx->stream);
    vc_free(sums);
    if (rc != VC_OK)
        return rc;
    if (e != hipSuccess || e2 != hipSuccess)
        return vc_fail(ctx, VC_ERR_HIP, "scan: %s",
                       hipGetErrorString(e != hipSuccess ? e : e2));
    return VC_OK;
}

// words of chunk totals scan_ws needs for n elements (all levels)
int64_t scan_tmp_words(int64_t n)
{
    int64_t t = 0;
    while (n > 1)
    {
        n = (n + 1023) / 1024;
        t += n;
    }
    return t > 0 ? t : 1;
}

// exclusive scan in place, queued on S without allocating or waiting; tmp
// holds scan_tmp_words(n) words
hipError_t scan_ws(hipStream_t S, uint32_t* a, int64_t n, uint32_t* tmp)
{
    const int64_t chunks = (n + 1023) / 1024;
    scan_chunks<<<(unsigned)chunks, 1024, 0, S>>>(a, n, tmp);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && chunks > 1)
    {
        e = scan_ws(S, tmp, chunks, tmp + chunks);
        if (e == hipSuccess)
        {
            scan_add<<<(unsigned)chunks, 1024, 0, S>>>(a, n, tmp);
            e = hipGetLastError();
        }
    }
    return e;
}

template <typename T>
int ws(vc_ctx* ctx, int slot, size_t bytes, T** p)
{
    return vc_ws_get(ctx, slot, bytes, reinterpret_cast<void**>(p));
}

// sites, quads and Euler terms of v into the volume's records.  Scratch
// comes from the context's slots; one wait for the totals (the Euler
// kernels are queued before it), the emit kernels are left queued.
int boundary_compute(vc_ctx* ctx, vc_vol* v, bool vec, int64_t etiles)
{
    hipStream_t S = ctx->stream;
    const Dims D{v->nx, v->ny, v->nz};
    const int64_t N = v->N;
    const int64_t ntiles = (N + kBlock - 1) / kBlock;
    const int64_t nch = vec ? N / kVz : 0;
    const size_t ncorner = (size_t)(v->nx + 1) * (v->ny + 1) * (v->nz + 1);
    uint32_t *ts, *tf, *cmap, *tmp, *ccnt = nullptr, *vmask = nullptr;
    unsigned long long* acc;
    uint8_t *bnd, *lay8 = nullptr;
    int32_t* lab;
    const int64_t tmpw = scan_tmp_words(etiles + 1);
    int rc;
    // tile totals carry one more word: the scan leaves the grand total there
    if ((rc = ws(ctx, VC_WS_BND_TS, 4 * (size_t)(etiles + 1), &ts)) != VC_OK ||
        (rc = ws(ctx, VC_WS_BND_TF, 4 * (size_t)(etiles + 1), &tf)) != VC_OK ||
        (rc = ws(ctx, VC_WS_BND_CMAP, 4 * ncorner, &cmap)) != VC_OK ||
        (rc = ws(ctx, VC_WS_SCAN, 4 * (size_t)tmpw, &tmp)) != VC_OK ||
        (rc = ws(ctx, VC_WS_BND_ACC, 8 * kSlots * kAcc + 4 * 256, &acc)) != VC_OK ||
        // boundary marks: one byte per 8-voxel chunk on the vectorised path,
        // one per voxel otherwise; labels: one per voxel (only those of
        // layer voxels are written and read)
        (rc = ws(ctx, VC_WS_BND_MARK, (size_t)(vec ? nch : N), &bnd)) != VC_OK ||
        (rc = ws(ctx, VC_WS_BND_LAB, 4 * (size_t)N, &lab)) != VC_OK)
        return rc;
    // per-chunk counts, per-voxel decision masks and thin-layer marks of the
    // vectorised passes
    if (vec && ((rc = ws(ctx, VC_WS_BND_CCNT, 4 * (size_t)nch, &ccnt)) != VC_OK ||
                (rc = ws(ctx, VC_WS_BND_VMASK, 4 * (size_t)N, &vmask)) != VC_OK ||
                (rc = ws(ctx, VC_WS_BND_LAY8, (size_t)nch, &lay8)) != VC_OK))
        return rc;
    int32_t* dtbl = reinterpret_cast<int32_t*>(acc + kSlots * kAcc);
    int32_t htbl[256];
    cube_table(htbl);

    hipError_t e = hipMemsetAsync(ts + etiles, 0, 4, S);
    if (e == hipSuccess)
        e = hipMemsetAsync(tf + etiles, 0, 4, S);
    if (e == hipSuccess)
        e = hipMemsetAsync(acc, 0, 8 * kSlots * kAcc, S);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dtbl, htbl, sizeof(htbl), hipMemcpyHostToDevice, S);
    // the Euler terms and the thin-layer components (vectorised path) need only
    // the occupancy: they run on the second stream, beside the counts, the
    // scans, the host's allocation of the outputs and the emit passes, and
    // join before the accumulators are read
    hipStream_t SE = vec ? ctx->stream_aux : S;
    if (e == hipSuccess && vec)
        e = hipEventRecord(ctx->ev_fork, S);
    if (e == hipSuccess && vec)
        e = hipStreamWaitEvent(SE, ctx->ev_fork, 0);
    if (e == hipSuccess && vec)
    {
        const int64_t nzc = v->nz / kVz;
        const int64_t lrows = (int64_t)(v->nx + 1) * (v->ny + 1) * nzc;
        const unsigned gl = (unsigned)((lrows + kBlock - 1) / kBlock);
        const unsigned gv = (unsigned)etiles;
        euler_rows_v<<<gl, kBlock, 0, SE>>>(v->occ, D, dtbl, acc);
        layer_boundary_v<<<gv, kBlock, 0, SE>>>(v->occ, D, bnd);
        const int tiles_x = (v->nx + kTX - 1) / kTX, tiles_y = (v->ny + kTY - 1) / kTY,
                  tiles_z = (int)((nzc + kTZ - 1) / kTZ);
        const int64_t n_tiles = (int64_t)tiles_x * tiles_y * tiles_z;
        layer_tile<<<(unsigned)(8 * ((n_tiles + 7) / 8)), kTile, 0, SE>>>(
            v->occ, bnd, D, tiles_y, tiles_z, n_tiles, lay8, lab, acc);
        layer_border<<<gv, kBlock, 0, SE>>>(lay8, D, lab, acc);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipEventRecord(ctx->ev_join, SE);
    }
    if (e == hipSuccess)
    {
        if (vec)
            bnd_count_v<<<(unsigned)etiles, kBlock, 0, S>>>(v->occ, D, ccnt, vmask, ts, tf);
        else
            bnd_count<<<(unsigned)etiles, kBlock, 0, S>>>(v->occ, D, N, ts, tf);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = scan_ws(S, ts, etiles + 1, tmp);
    if (e == hipSuccess)
        e = scan_ws(S, tf, etiles + 1, tmp);
    uint32_t tot[2] = {0, 0};
    if (e == hipSuccess)
        e = hipMemcpyAsync(&tot[0], ts + etiles, 4, hipMemcpyDeviceToHost, S);
    if (e == hipSuccess)
        e = hipMemcpyAsync(&tot[1], tf + etiles, 4, hipMemcpyDeviceToHost, S);
    // Euler terms (the per-voxel path: on this stream)
    if (e == hipSuccess && !vec)
    {
        const dim3 gl((unsigned)((v->nz + 1 + kBlock - 1) / kBlock),
                      (unsigned)(v->ny + 1), (unsigned)(v->nx + 1));
        euler_lattice<<<gl, kBlock, 0, S>>>(v->occ, D, dtbl, acc);
        chi_lattice<<<gl, kBlock, 0, S>>>(v->occ, D, acc);
        const dim3 gv((unsigned)((v->nz + kBlock - 1) / kBlock),
                      (unsigned)v->ny, (unsigned)v->nx);
        euler_voxels<<<gv, kBlock, 0, S>>>(v->occ, D, acc);
        layer_boundary<<<gv, kBlock, 0, S>>>(v->occ, D, bnd);
        layer_init<<<gv, kBlock, 0, S>>>(v->occ, bnd, D, lab);
        layer_union<<<gv, kBlock, 0, S>>>(lab, D, acc + 7);
        layer_count<<<(unsigned)ntiles, kBlock, 0, S>>>(lab, N, acc);
        e = hipGetLastError();
    }
    unsigned long long hacc[kSlots * kAcc];
    if (e == hipSuccess && !vec)
        e = hipMemcpyAsync(hacc, acc, sizeof(hacc), hipMemcpyDeviceToHost, S);
    if (e == hipSuccess)
        e = hipStreamSynchronize(S);
    if (e != hipSuccess)
    {
        if (vec)
            (void)hipStreamSynchronize(SE);
        return vc_fail(ctx, VC_ERR_HIP, "boundary count/Euler: %s", hipGetErrorString(e));
    }
    const int64_t ns = tot[0], nf = tot[1];

    vc_free(v->bnd_sites);
    vc_free(v->bnd_quads);
    v->bnd_sites = nullptr;
    v->bnd_quads = nullptr;
    if ((rc = vc_alloc(ctx, (void**)&v->bnd_sites, 12 * (size_t)ns)) != VC_OK ||
        (rc = vc_alloc(ctx, (void**)&v->bnd_quads, 16 * (size_t)nf)) != VC_OK)
    {
        if (vec)
            (void)hipStreamSynchronize(SE); // the workspace's last user is done
        return rc;
    }
    if (vec)
    {
        bnd_emit_sites_v<<<(unsigned)etiles, kBlock, 0, S>>>(D, ccnt, vmask, ts,
                                                             v->bnd_sites, cmap);
        bnd_emit_quads_v<<<(unsigned)etiles, kBlock, 0, S>>>(D, ccnt, vmask, tf, cmap,
                                                             v->bnd_quads);
    }
    else
    {
        bnd_emit_sites<<<(unsigned)etiles, kBlock, 0, S>>>(v->occ, D, N, ts,
                                                           v->bnd_sites, cmap);
        bnd_emit_quads<<<(unsigned)etiles, kBlock, 0, S>>>(v->occ, D, N, tf, cmap,
                                                           v->bnd_quads);
    }
    e = hipGetLastError();
    // join the Euler chain, then read the accumulators
    if (e == hipSuccess && vec)
        e = hipStreamWaitEvent(S, ctx->ev_join, 0);
    if (e == hipSuccess && vec)
        e = hipMemcpyAsync(hacc, acc, sizeof(hacc), hipMemcpyDeviceToHost, S);
    if (e == hipSuccess && vec)
        e = hipStreamSynchronize(S);
    if (e != hipSuccess)
    {
        if (vec)
            (void)hipStreamSynchronize(SE);
        return vc_fail(ctx, VC_ERR_HIP, "boundary emit: %s", hipGetErrorString(e));
    }
    long long sum[kAcc] = {0};
    for (int sl = 0; sl < kSlots; ++sl)
        for (int q = 0; q < kAcc; ++q)
            sum[q] += (long long)hacc[sl * kAcc + q];
    if (sum[7] != 0)
        return vc_fail(ctx, VC_ERR_GUARD, "boundary: union-find retry bound reached");

    vc_euler& E = v->bnd_euler;
    std::memset(&E, 0, sizeof(E));
    E.V_bnd = sum[0];
    E.E_bnd = sum[1];
    E.V_int = sum[2] / 8;
    E.E_int = sum[3];
    E.F_int = sum[4] / 2;
    E.T = sum[5];
    // vectorised path: tile-local roots - cross-tile links (slot 10 is 0
    // otherwise)
    E.C = sum[6] - sum[10];
    E.V = E.V_bnd + E.V_int;
    E.E = E.E_bnd + E.E_int;
    E.F = nf + E.F_int;
    E.euler = (int64_t)(int32_t)(E.V - E.E + E.F - E.T);
    E.chi = sum[8] - sum[9];
    v->bnd_nsites = ns;
    v->bnd_nquads = nf;
    v->bnd_done = true;
    return VC_OK;
}

} // namespace

int vc_scan_u32(vc_ctx* ctx, uint32_t* a, int64_t n)
{
    return scan_inplace(ctx, a, n);
}

int vc_boundary_run(vc_ctx* ctx, vc_vol* v, float* sites, int64_t site_cap,
                    int64_t* n_sites, int32_t* quads, int64_t quad_cap,
                    int64_t* n_quads, vc_euler* eu)
{
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    const int64_t N = v->N;
    const int64_t ntiles = (N + kBlock - 1) / kBlock;
    // site and face counts are scanned in 32 bits and quads index sites in
    // int32: a volume has at most 3N + 2(nx ny + ny nz + nx nz) boundary faces
    // and (nx+1)(ny+1)(nz+1) corners
    {
        const int64_t faces_max =
            3 * N + 2 * ((int64_t)v->nx * v->ny + (int64_t)v->ny * v->nz +
                         (int64_t)v->nx * v->nz);
        const int64_t corners =
            (int64_t)(v->nx + 1) * (v->ny + 1) * (v->nz + 1);
        if (faces_max > (int64_t)UINT32_MAX || corners > (int64_t)INT32_MAX)
            return vc_fail(ctx, VC_ERR_INVALID,
                           "boundary: %d x %d x %d may have more faces than 32-bit "
                           "counts or more corners than int32 site indices hold",
                           v->nx, v->ny, v->nz);
    }
    // site / face tiles: 256 voxels, or 256 chunks of 8 z when nz allows
    const bool vec = (v->nz % kVz) == 0;
    const int64_t etiles =
        vec ? (N / kVz + kBlock - 1) / kBlock : ntiles;
    if (!v->bnd_done)
    {
        const int rc = boundary_compute(ctx, v, vec, etiles);
        if (rc != VC_OK)
            return rc;
    }
    *n_sites = v->bnd_nsites;
    *n_quads = v->bnd_nquads;
    if (eu)
        *eu = v->bnd_euler;
    if ((sites && site_cap < v->bnd_nsites) || (quads && quad_cap < v->bnd_nquads))
        return vc_fail(ctx, VC_ERR_CAPACITY,
                       "boundary: need %lld sites and %lld quads",
                       (long long)v->bnd_nsites, (long long)v->bnd_nquads);
    if (sites && v->bnd_nsites)
        VC_HIP(ctx, hipMemcpyAsync(sites, v->bnd_sites, 12 * (size_t)v->bnd_nsites,
                                   hipMemcpyDeviceToHost, S));
    if (quads && v->bnd_nquads)
        VC_HIP(ctx, hipMemcpyAsync(quads, v->bnd_quads, 16 * (size_t)v->bnd_nquads,
                                   hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    return VC_OK;
}
