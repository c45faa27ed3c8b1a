// vc_voro.hip -- inside tagging of points and lambda measures of a 2-complex
// on the device (include/vc_voro.h), gfx950.
//
// Both are gathers: one thread per point / face / edge / vertex, the measure
// of an edge or vertex a maximum over its CSR neighbourhood.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <vector>

#include "vc_internal.h"
#include "../../include/vc_voro.h"

namespace
{

constexpr int kBlock = 256;

// std::round (half away from zero) then the occupancy of that voxel; points
// whose rounded coordinate does not fit an int are outside
__global__ void __launch_bounds__(kBlock)
    tag_points(const float* __restrict__ pts, int64_t n,
               const uint8_t* __restrict__ occ, int nx, int ny, int nz,
               uint8_t* __restrict__ valid)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int v[3];
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        const float r = roundf(pts[3 * i + a]);
        ok = ok && (r > -1.0e9f && r < 1.0e9f); // also false for NaN
        v[a] = ok ? (int)r : 0;
    }
    ok = ok && (unsigned)v[0] < (unsigned)nx && (unsigned)v[1] < (unsigned)ny &&
         (unsigned)v[2] < (unsigned)nz;
    valid[i] = ok && occ[((int64_t)v[0] * ny + v[1]) * nz + v[2]] != 0;
}

// |a - b| as trimesh::dist: ((d0^2 + d1^2) + d2^2), no contraction,
// correctly rounded square root
__device__ float site_dist(const float* __restrict__ s, int a, int b)
{
#pragma clang fp contract(off)
    const float d0 = s[3 * b] - s[3 * a];
    const float d1 = s[3 * b + 1] - s[3 * a + 1];
    const float d2 = s[3 * b + 2] - s[3 * a + 2];
    const float q = (d0 * d0 + d1 * d1) + d2 * d2;
    return sqrtf(q);
}

__global__ void __launch_bounds__(kBlock)
    face_lambda(const float* __restrict__ sites,
                const int32_t* __restrict__ face_sites, int32_t nf,
                float* __restrict__ lam_f)
{
    const int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf)
        return;
    lam_f[f] = site_dist(sites, face_sites[2 * f], face_sites[2 * f + 1]);
}

// OffT: int64_t offsets (host arrays) or uint32_t (a resident vc_cc)
template <typename OffT>
__global__ void __launch_bounds__(kBlock)
    edge_lambda(const float* __restrict__ lam_f,
                const uint8_t* __restrict__ face_valid,
                const OffT* __restrict__ ef_off,
                const int32_t* __restrict__ ef, int32_t ne,
                float* __restrict__ lam_e)
{
    const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne)
        return;
    float m = 0.0f;
    for (int64_t j = (int64_t)ef_off[e]; j < (int64_t)ef_off[e + 1]; ++j)
    {
        const int32_t f = ef[j];
        if (face_valid[f])
            m = fmaxf(m, lam_f[f]);
    }
    lam_e[e] = m;
}

template <typename OffT>
__global__ void __launch_bounds__(kBlock)
    vertex_lambda(const float* __restrict__ lam_f,
                  const uint8_t* __restrict__ face_valid,
                  const OffT* __restrict__ ve_off,
                  const int32_t* __restrict__ ve,
                  const OffT* __restrict__ ef_off,
                  const int32_t* __restrict__ ef, int32_t nv,
                  float* __restrict__ lam_v)
{
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv)
        return;
    float m = 0.0f;
    for (int64_t i = (int64_t)ve_off[v]; i < (int64_t)ve_off[v + 1]; ++i)
    {
        const int32_t e = ve[i];
        for (int64_t j = (int64_t)ef_off[e]; j < (int64_t)ef_off[e + 1]; ++j)
        {
            const int32_t f = ef[j];
            if (face_valid[f])
                m = fmaxf(m, lam_f[f]);
        }
    }
    lam_v[v] = m;
}

// last[v] = the highest face id whose polygon holds v (last starts at -1)
__global__ void __launch_bounds__(kBlock)
    face_last(const int64_t* __restrict__ face_off,
              const int32_t* __restrict__ face_vts, int32_t nf,
              int32_t* __restrict__ last)
{
    const int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf)
        return;
    for (int64_t j = face_off[f]; j < face_off[f + 1]; ++j)
        atomicMax(&last[face_vts[j]], f);
}

// r[v] = |v - first site of face last[v]| (trimesh::dist order), 0 if v is
// in no face
__global__ void __launch_bounds__(kBlock)
    vertex_radius(const float* __restrict__ vts, const float* __restrict__ sites,
                  const int32_t* __restrict__ face_sites,
                  const int32_t* __restrict__ last, int32_t nv,
                  float* __restrict__ r)
{
#pragma clang fp contract(off)
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv)
        return;
    const int32_t f = last[v];
    if (f < 0)
    {
        r[v] = 0.0f;
        return;
    }
    const int32_t s = face_sites[2 * f];
    const float d0 = vts[3 * v] - sites[3 * s];
    const float d1 = vts[3 * v + 1] - sites[3 * s + 1];
    const float d2 = vts[3 * v + 2] - sites[3 * s + 2];
    r[v] = sqrtf((d0 * d0 + d1 * d1) + d2 * d2);
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// device copies that are freed on every exit path
struct Scratch
{
    std::vector<void*> p;
    ~Scratch()
    {
        for (void* q : p)
            vc_free(q);
    }
    template <typename T>
    int get(vc_ctx* ctx, T** out, size_t count)
    {
        void* q = nullptr;
        int rc = vc_alloc(ctx, &q, count * sizeof(T) + 16);
        if (rc == VC_OK)
            p.push_back(q);
        *out = static_cast<T*>(q);
        return rc;
    }
};

// ---- discrete Voronoi 2-complex from the feature transform
// A face for every pair of 6-adjacent inside voxels p, q = p + e_a whose
// nearest background voxels differ; faces ordered by p (linear), then axis
// x, y, z.  The face lies at lattice coordinate p_a + 1 along axis a; its
// corners are the lattice points (p_b + db, p_c + dc) in the cyclic order
// (0,0), (1,0), (1,1), (0,1) of the two other axes b = a+1, c = a+2 (mod 3),
// so every face is counter-clockwise seen from +a.  Lattice point (i, j, k)
// is the voxel corner at (i - 1/2, j - 1/2, k - 1/2).
struct FtGrid
{
    const uint8_t* occ;
    const uint32_t* ft;
    int nx, ny, nz;
    vc_pack pk;
};

__device__ __forceinline__ uint32_t vft_axes(const FtGrid& G, int x, int y,
                                             int z, int64_t i)
{
    if (!G.occ[i])
        return 0u;
    const uint32_t f = G.ft[i];
    uint32_t m = 0;
    const int64_t sy = G.nz, sx = (int64_t)G.ny * G.nz;
    if (x + 1 < G.nx && G.occ[i + sx] && G.ft[i + sx] != f)
        m |= 1u;
    if (y + 1 < G.ny && G.occ[i + sy] && G.ft[i + sy] != f)
        m |= 2u;
    if (z + 1 < G.nz && G.occ[i + 1] && G.ft[i + 1] != f)
        m |= 4u;
    return m;
}

__device__ __forceinline__ int64_t vft_site(const FtGrid& G, uint32_t f)
{
    int x, y, z;
    vc_unpack_xyz(G.pk, f, x, y, z);
    return ((int64_t)x * G.ny + y) * G.nz + z;
}

// corner lattice index of corner q (0..3) of the face of voxel (x, y, z)
// towards +a
__device__ __forceinline__ int64_t vft_corner(const FtGrid& G, int x, int y,
                                              int z, int a, int q)
{
    int c[3] = {x, y, z};
    c[a] += 1;
    c[(a + 1) % 3] += (q == 1 || q == 2) ? 1 : 0;
    c[(a + 2) % 3] += (q >= 2) ? 1 : 0;
    return ((int64_t)c[0] * (G.ny + 1) + c[1]) * (G.nz + 1) + c[2];
}

__device__ __forceinline__ void vft_xyz(const FtGrid& G, int64_t i, int& x,
                                        int& y, int& z)
{
    z = (int)(i % G.nz);
    const int64_t r = i / G.nz;
    y = (int)(r % G.ny);
    x = (int)(r / G.ny);
}

// pass 1: faces per 256-voxel tile; flags the used lattice points and sites
__global__ void __launch_bounds__(kBlock)
    vft_count(FtGrid G, int64_t N, uint32_t* __restrict__ tile_n,
              uint32_t* __restrict__ lat_flag, uint32_t* __restrict__ site_flag)
{
    __shared__ uint32_t red[kBlock / 64];
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t m = 0;
    if (i < N)
    {
        int x, y, z;
        vft_xyz(G, i, x, y, z);
        m = vft_axes(G, x, y, z, i);
        if (m)
        {
            site_flag[vft_site(G, G.ft[i])] = 1u;
            const int64_t sy = G.nz, sx = (int64_t)G.ny * G.nz;
            const int64_t nb[3] = {i + sx, i + sy, i + 1};
            for (int a = 0; a < 3; ++a)
            {
                if (!((m >> a) & 1u))
                    continue;
                site_flag[vft_site(G, G.ft[nb[a]])] = 1u;
                for (int q = 0; q < 4; ++q)
                    lat_flag[vft_corner(G, x, y, z, a, q)] = 1u;
            }
        }
    }
    uint32_t c = (uint32_t)__popc(m);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        c += __shfl_xor(c, d);
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        uint32_t t = 0;
        for (int w = 0; w < kBlock / 64; ++w)
            t += red[w];
        tile_n[blockIdx.x] = t;
    }
}

// pass 2: faces (4 lattice ids through the scanned lattice flags) and their
// two site ids (scanned site flags), at the tile offset + the voxel's rank
__global__ void __launch_bounds__(kBlock)
    vft_emit(FtGrid G, int64_t N, const uint32_t* __restrict__ tile_off,
             const uint32_t* __restrict__ lat_id,
             const uint32_t* __restrict__ site_id, int32_t* __restrict__ faces,
             int32_t* __restrict__ face_sites)
{
    __shared__ uint32_t wsum[kBlock / 64];
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t m = 0;
    int x = 0, y = 0, z = 0;
    if (i < N)
    {
        vft_xyz(G, i, x, y, z);
        m = vft_axes(G, x, y, z, i);
    }
    // exclusive prefix of the per-voxel counts (0..3) over the tile, from
    // one ballot per bit plane
    const uint32_t c = (uint32_t)__popc(m);
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int bit = 0; bit < 2; ++bit)
    {
        const uint64_t b = __ballot((c >> bit) & 1u);
        pre += __builtin_amdgcn_mbcnt_hi(
                   (uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u))
               << bit;
        tot += (uint32_t)__popcll(b) << bit;
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        wsum[wv] = tot;
    __syncthreads();
    for (int w = 0; w < wv; ++w)
        pre += wsum[w];
    if (!m)
        return;
    int64_t f = (int64_t)tile_off[blockIdx.x] + pre;
    const int64_t sy = G.nz, sx = (int64_t)G.ny * G.nz;
    const int64_t nb[3] = {i + sx, i + sy, i + 1};
    const int32_t s0 = (int32_t)site_id[vft_site(G, G.ft[i])];
    for (int a = 0; a < 3; ++a)
    {
        if (!((m >> a) & 1u))
            continue;
        for (int q = 0; q < 4; ++q)
            faces[4 * f + q] = (int32_t)lat_id[vft_corner(G, x, y, z, a, q)];
        face_sites[2 * f] = s0;
        face_sites[2 * f + 1] = (int32_t)site_id[vft_site(G, G.ft[nb[a]])];
        ++f;
    }
}

// pass 3: coordinates of the used lattice points (scanned flags: point l is
// used iff id[l + 1] - id[l] == 1)
__global__ void __launch_bounds__(kBlock)
    vft_vertices(int nx, int ny, int nz, const uint32_t* __restrict__ lat_id,
                 float* __restrict__ vts)
{
    const int64_t L = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    const int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (l >= L || lat_id[l + 1] == lat_id[l])
        return;
    const int k = (int)(l % (nz + 1));
    const int64_t r = l / (nz + 1);
    const int j = (int)(r % (ny + 1)), ii = (int)(r / (ny + 1));
    const int64_t v = lat_id[l];
    vts[3 * v] = (float)ii - 0.5f;
    vts[3 * v + 1] = (float)j - 0.5f;
    vts[3 * v + 2] = (float)k - 0.5f;
}

__global__ void __launch_bounds__(kBlock)
    vft_sites(FtGrid G, int64_t N, const uint32_t* __restrict__ site_id,
              float* __restrict__ sites)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= N || site_id[i + 1] == site_id[i])
        return;
    int x, y, z;
    vft_xyz(G, i, x, y, z);
    const int64_t s = site_id[i];
    sites[3 * s] = (float)x;
    sites[3 * s + 1] = (float)y;
    sites[3 * s + 2] = (float)z;
}

// ---- inside extraction on the device (the order of vc_voro_extract_inside)
constexpr uint64_t kEmptyKey = ~0ull;

__device__ __forceinline__ uint64_t pair_key(int32_t a, int32_t b)
{
    return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}

__device__ __forceinline__ uint64_t mix64(uint64_t k)
{
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    return k;
}

// insert key with value v (the table keeps the smallest value per key)
__device__ void hash_min(uint64_t* keys, unsigned long long* vals, uint64_t mask,
                         uint64_t key, unsigned long long v)
{
    uint64_t h = mix64(key) & mask;
    for (;;)
    {
        const uint64_t prev = atomicCAS(
            reinterpret_cast<unsigned long long*>(&keys[h]),
            (unsigned long long)kEmptyKey, (unsigned long long)key);
        if (prev == kEmptyKey || prev == key)
        {
            atomicMin(&vals[h], v);
            return;
        }
        h = (h + 1) & mask;
    }
}

__device__ unsigned long long hash_get(const uint64_t* keys,
                                       const unsigned long long* vals,
                                       uint64_t mask, uint64_t key)
{
    uint64_t h = mix64(key) & mask;
    while (keys[h] != key)
        h = (h + 1) & mask;
    return vals[h];
}

__global__ void __launch_bounds__(kBlock)
    xi_flags(int32_t nv, int32_t ne, int32_t nf, const int32_t* __restrict__ E,
             const int64_t* __restrict__ fo, const int32_t* __restrict__ fv,
             const uint8_t* __restrict__ valid, uint32_t* __restrict__ vflag,
             uint32_t* __restrict__ eflag, uint32_t* __restrict__ tcount)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < nv)
        vflag[i] = valid[i] ? 1u : 0u;
    if (i < ne)
        eflag[i] = valid[E[2 * i]] && valid[E[2 * i + 1]] ? 1u : 0u;
    if (i < nf)
    {
        bool ok = true;
        for (int64_t j = fo[i]; j < fo[i + 1] && ok; ++j)
            ok = valid[fv[j]] != 0;
        const int64_t n = fo[i + 1] - fo[i];
        tcount[i] = ok && n >= 3 ? (uint32_t)(n - 2) : 0u;
    }
}

// valid vertex list, valid edges (renumbered, measures) and their exact
// pairs in the table with value 0
__global__ void __launch_bounds__(kBlock)
    xi_emit_ve(int32_t nv, int32_t ne, const int32_t* __restrict__ E,
               const uint8_t* __restrict__ valid,
               const uint32_t* __restrict__ vpos, const uint32_t* __restrict__ epos,
               const float* __restrict__ lam_e, int32_t* __restrict__ vts_idx,
               int32_t* __restrict__ e_out, float* __restrict__ em,
               uint64_t* keys, unsigned long long* vals, uint64_t mask)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < nv && valid[i])
        vts_idx[vpos[i]] = (int32_t)i;
    if (i < ne && epos[i + 1] != epos[i])
    {
        const int32_t a = E[2 * i], b = E[2 * i + 1];
        const uint32_t k = epos[i];
        e_out[2 * k] = (int32_t)vpos[a];
        e_out[2 * k + 1] = (int32_t)vpos[b];
        em[k] = lam_e[i];
        hash_min(keys, vals, mask, pair_key(a, b), 0ull);
    }
}

// triangles of the valid faces (zig-zag: triangle t of an n-gon is
// (p[i], p[j], p[j+1]) for even t and (p[i], p[j], p[i-1]) for odd t, with
// i = n-1-floor(t/2), j = ceil(t/2)); their normalised sides as candidates
// 3 t + s, inserted with value candidate + 1
__global__ void __launch_bounds__(kBlock)
    xi_tris(int32_t nf, const int64_t* __restrict__ fo,
            const int32_t* __restrict__ fv, const uint32_t* __restrict__ toff,
            const uint32_t* __restrict__ vpos, const float* __restrict__ lam_f,
            int32_t* __restrict__ tris, int32_t* __restrict__ from_fi,
            float* __restrict__ fm, uint64_t* __restrict__ ckey, uint64_t* keys,
            unsigned long long* vals, uint64_t mask)
{
    const int64_t f = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (f >= nf || toff[f + 1] == toff[f])
        return;
    const int32_t* p = fv + fo[f];
    const int n = (int)(fo[f + 1] - fo[f]);
    for (int t = 0; t < n - 2; ++t)
    {
        const int i = n - 1 - t / 2, j = (t + 1) / 2;
        const int32_t v[3] = {p[i], p[j], (t & 1) ? p[i - 1] : p[j + 1]};
        const int64_t T = (int64_t)toff[f] + t;
        for (int s = 0; s < 3; ++s)
        {
            tris[3 * T + s] = (int32_t)vpos[v[s]];
            int32_t a = v[s], b = v[(s + 1) % 3];
            if (a > b)
            {
                const int32_t x = a;
                a = b;
                b = x;
            }
            const uint64_t k = pair_key(a, b);
            ckey[3 * T + s] = k;
            hash_min(keys, vals, mask, k, (unsigned long long)(3 * T + s) + 1ull);
        }
        from_fi[T] = (int32_t)f;
        fm[T] = lam_f[f];
    }
}

__global__ void __launch_bounds__(kBlock)
    xi_keep(int64_t nc, const uint64_t* __restrict__ ckey, const uint64_t* keys,
            const unsigned long long* vals, uint64_t mask,
            uint32_t* __restrict__ keep)
{
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c < nc)
        keep[c] = hash_get(keys, vals, mask, ckey[c]) == (unsigned long long)c + 1ull
                      ? 1u
                      : 0u;
}

__global__ void __launch_bounds__(kBlock)
    xi_emit_sides(int64_t nc, const uint64_t* __restrict__ ckey,
                  const uint32_t* __restrict__ kpos,
                  const uint32_t* __restrict__ vpos,
                  const int32_t* __restrict__ from_fi,
                  const float* __restrict__ lam_f, int64_t base,
                  int32_t* __restrict__ e_out, float* __restrict__ em)
{
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= nc || kpos[c + 1] == kpos[c])
        return;
    const int64_t k = base + kpos[c];
    const uint64_t key = ckey[c];
    e_out[2 * k] = (int32_t)vpos[(int32_t)(key >> 32)];
    e_out[2 * k + 1] = (int32_t)vpos[(int32_t)(key & 0xFFFFFFFFull)];
    em[k] = lam_f[from_fi[c / 3]];
}

// ---- cellcomplex::finalize on the device: candidates are the given edges
// (exact pairs) then the face sides (normalised), in that order; the first
// candidate of each pair becomes an edge, numbered in candidate order
__global__ void __launch_bounds__(kBlock)
    fz_keys(int32_t ne, const int32_t* __restrict__ E, int32_t nf,
            const int64_t* __restrict__ fo, const int32_t* __restrict__ fv,
            uint64_t* __restrict__ ckey, uint64_t* keys,
            unsigned long long* vals, uint64_t mask)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < ne)
    {
        const uint64_t k = pair_key(E[2 * i], E[2 * i + 1]);
        ckey[i] = k;
        hash_min(keys, vals, mask, k, (unsigned long long)i);
    }
    if (i < nf)
    {
        const int64_t n = fo[i + 1] - fo[i];
        for (int64_t j = 0; j < n; ++j)
        {
            int32_t a = fv[fo[i] + j], b = fv[fo[i] + (j + 1) % n];
            if (a > b)
            {
                const int32_t t = a;
                a = b;
                b = t;
            }
            const uint64_t k = pair_key(a, b);
            const int64_t c = ne + fo[i] + j;
            ckey[c] = k;
            hash_min(keys, vals, mask, k, (unsigned long long)c);
        }
    }
}

__global__ void __launch_bounds__(kBlock)
    fz_first(int64_t nc, const uint64_t* __restrict__ ckey, const uint64_t* keys,
             const unsigned long long* vals, uint64_t mask,
             uint32_t* __restrict__ first)
{
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c < nc)
        first[c] = hash_get(keys, vals, mask, ckey[c]) == (unsigned long long)c ? 1u
                                                                                : 0u;
}

__global__ void __launch_bounds__(kBlock)
    fz_emit(int64_t nc, int32_t ne, const uint64_t* __restrict__ ckey,
            const uint32_t* __restrict__ pos, const uint64_t* keys,
            const unsigned long long* vals, uint64_t mask,
            int32_t* __restrict__ e_out, int32_t* __restrict__ fe)
{
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= nc)
        return;
    const uint64_t k = ckey[c];
    if (pos[c + 1] != pos[c])
    {
        e_out[2 * (int64_t)pos[c]] = (int32_t)(k >> 32);
        e_out[2 * (int64_t)pos[c] + 1] = (int32_t)(k & 0xFFFFFFFFull);
    }
    if (c >= ne)
        fe[c - ne] = (int32_t)pos[hash_get(keys, vals, mask, k)];
}

__global__ void __launch_bounds__(kBlock) fill_u64(uint64_t* a, int64_t n, uint64_t v)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n)
        a[i] = v;
}

// counting pass of the discrete diagram: faces per tile, used lattice points
// and sites flagged, all three scanned; tot = faces, vertices, sites
int ft_counts(vc_ctx* ctx, vc_vol* vol, Scratch& sc, uint32_t** tile_out,
              uint32_t** lat_out, uint32_t** sid_out, int64_t tot[3])
{
    const int64_t N = vol->N;
    const int64_t L = (int64_t)(vol->nx + 1) * (vol->ny + 1) * (vol->nz + 1);
    // face counts are scanned in uint32 (at most 3 per voxel), vertex ids are
    // int32
    if (3 * N >= (int64_t)UINT32_MAX || L >= (int64_t)INT32_MAX)
        return vc_fail(ctx, VC_ERR_INVALID,
                       "vc_voro_from_ft: volume too large (%lld voxels)",
                       (long long)N);
    hipStream_t S = ctx->stream;
    const int64_t tiles = (N + kBlock - 1) / kBlock;
    uint32_t *tile = nullptr, *lat = nullptr, *sid = nullptr;
    int rc = sc.get(ctx, &tile, (size_t)tiles + 1);
    if (rc == VC_OK)
        rc = sc.get(ctx, &lat, (size_t)L + 1);
    if (rc == VC_OK)
        rc = sc.get(ctx, &sid, (size_t)N + 1);
    if (rc != VC_OK)
        return rc;
    VC_HIP(ctx, hipMemsetAsync(tile, 0, 4 * ((size_t)tiles + 1), S));
    VC_HIP(ctx, hipMemsetAsync(lat, 0, 4 * ((size_t)L + 1), S));
    VC_HIP(ctx, hipMemsetAsync(sid, 0, 4 * ((size_t)N + 1), S));
    const FtGrid G{vol->occ, vol->ft, vol->nx, vol->ny, vol->nz, vol->pk};
    vft_count<<<(unsigned)tiles, kBlock, 0, S>>>(G, N, tile, lat, sid);
    VC_CHECK_LAUNCH(ctx);
    if ((rc = vc_scan_u32(ctx, tile, tiles + 1)) != VC_OK ||
        (rc = vc_scan_u32(ctx, lat, L + 1)) != VC_OK ||
        (rc = vc_scan_u32(ctx, sid, N + 1)) != VC_OK)
        return rc;
    uint32_t t[3];
    VC_HIP(ctx, hipMemcpy(&t[0], tile + tiles, 4, hipMemcpyDeviceToHost));
    VC_HIP(ctx, hipMemcpy(&t[1], lat + L, 4, hipMemcpyDeviceToHost));
    VC_HIP(ctx, hipMemcpy(&t[2], sid + N, 4, hipMemcpyDeviceToHost));
    tot[0] = t[0];
    tot[1] = t[1];
    tot[2] = t[2];
    if (tot[0] > (int64_t)INT32_MAX)
        return vc_fail(ctx, VC_ERR_INVALID,
                       "vc_voro_from_ft: %lld faces exceed int32 ids",
                       (long long)tot[0]);
    *tile_out = tile;
    *lat_out = lat;
    *sid_out = sid;
    return VC_OK;
}

template <typename T>
int own_alloc(vc_ctx* ctx, T** p, size_t n)
{
    void* q = nullptr;
    const int rc = vc_alloc(ctx, &q, n * sizeof(T) + 16);
    *p = static_cast<T*>(q);
    return rc;
}

} // namespace

// host-side validation of a complex given as arrays (ids in range,
// monotone face offsets starting at 0, total size within int32)
int vc_cc_check_host(vc_ctx* ctx, int32_t nv, int32_t ne_in,
                     const int32_t* edges_in, int32_t nf,
                     const int64_t* face_off, const int32_t* face_vts)
{
    const int64_t nfv = nf ? face_off[nf] : 0;
    if ((nf && face_off[0] != 0) || nfv < 0 || (int64_t)ne_in + nfv >= (int64_t)INT32_MAX)
        return vc_fail(ctx, VC_ERR_INVALID, "complex: bad face offsets or too large");
    for (int32_t f = 0; f < nf; ++f)
        if (face_off[f + 1] < face_off[f])
            return vc_fail(ctx, VC_ERR_INVALID, "complex: face offsets decrease");
    for (int64_t i = 0; i < 2LL * ne_in; ++i)
        if (edges_in[i] < 0 || edges_in[i] >= nv)
            return vc_fail(ctx, VC_ERR_INVALID, "complex: edge end out of range");
    for (int64_t j = 0; j < nfv; ++j)
        if (face_vts[j] < 0 || face_vts[j] >= nv)
            return vc_fail(ctx, VC_ERR_INVALID, "complex: face vertex out of range");
    return VC_OK;
}

// cellcomplex::finalize on device arrays: deo receives the edges (capacity
// 2 * (ne_in + nfv) ints), dfe each face side's edge id (nfv ints); *ne_out
// the edge count.  Waits for the count.
int vc_cc_finalize_dev(vc_ctx* ctx, int32_t ne_in, const int32_t* dE, int32_t nf,
                       const int64_t* dfo, const int32_t* dfv, int64_t nfv,
                       int32_t* deo, int32_t* dfe, int32_t* ne_out)
{
    hipStream_t S = ctx->stream;
    Scratch sc;
    const int64_t NC = (int64_t)ne_in + nfv;
    uint64_t H = 1024;
    while (H < 2 * (uint64_t)NC + 1)
        H <<= 1;
    uint64_t *keys, *ckey;
    unsigned long long* vals;
    uint32_t* pos;
    int rc = VC_OK;
    auto get = [&](auto** p, size_t n) {
        if (rc == VC_OK)
            rc = sc.get(ctx, p, n);
    };
    get(&keys, (size_t)H);
    get(&vals, (size_t)H);
    get(&ckey, (size_t)NC + 1);
    get(&pos, (size_t)NC + 1);
    if (rc != VC_OK)
        return rc;
    fill_u64<<<blocks_for((int64_t)H), kBlock, 0, S>>>(keys, (int64_t)H, kEmptyKey);
    fill_u64<<<blocks_for((int64_t)H), kBlock, 0, S>>>(
        reinterpret_cast<uint64_t*>(vals), (int64_t)H, ~0ull);
    VC_HIP(ctx, hipMemsetAsync(pos, 0, 4 * ((size_t)NC + 1), S));
    const int64_t mx = std::max<int64_t>(ne_in, nf);
    if (mx)
        fz_keys<<<blocks_for(mx), kBlock, 0, S>>>(ne_in, dE, nf, dfo, dfv, ckey, keys,
                                                  vals, H - 1);
    if (NC)
        fz_first<<<blocks_for(NC), kBlock, 0, S>>>(NC, ckey, keys, vals, H - 1, pos);
    VC_CHECK_LAUNCH(ctx);
    if ((rc = vc_scan_u32(ctx, pos, NC + 1)) != VC_OK)
        return rc;
    uint32_t ne = 0;
    VC_HIP(ctx, hipMemcpyAsync(&ne, pos + NC, 4, hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    *ne_out = (int32_t)ne;
    if (NC)
        fz_emit<<<blocks_for(NC), kBlock, 0, S>>>(NC, ne_in, ckey, pos, keys, vals,
                                                  H - 1, deo, dfe);
    VC_CHECK_LAUNCH(ctx);
    // the temporaries are freed on return: wait for the kernels reading them
    VC_HIP(ctx, hipStreamSynchronize(S));
    return VC_OK;
}

extern "C" {

int vc_voro_tag_points(vc_ctx* ctx, vc_vol* vol, const float* pts, int64_t n,
                       uint8_t* valid)
{
    if (!ctx || !vol || vol->ctx != ctx || n < 0 || (n > 0 && (!pts || !valid)))
        return VC_ERR_INVALID;
    if (n == 0)
        return VC_OK;
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    Scratch sc;
    float* dp;
    uint8_t* dv;
    int rc = sc.get(ctx, &dp, 3 * (size_t)n);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dv, (size_t)n);
    if (rc != VC_OK)
        return rc;
    VC_HIP(ctx, hipMemcpyAsync(dp, pts, 12 * (size_t)n, hipMemcpyHostToDevice,
                               S));
    tag_points<<<blocks_for(n), kBlock, 0, S>>>(dp, n, vol->occ, vol->nx,
                                                 vol->ny, vol->nz, dv);
    VC_CHECK_LAUNCH(ctx);
    VC_HIP(ctx, hipMemcpyAsync(valid, dv, (size_t)n, hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    return VC_OK;
}

int vc_voro_measures(vc_ctx* ctx, int32_t nv, int32_t ne, int32_t nf,
                     const float* sites, int32_t nsites,
                     const int32_t* face_sites, const uint8_t* face_valid,
                     const int64_t* ve_off, const int32_t* ve,
                     const int64_t* ef_off, const int32_t* ef, float* lam_v,
                     float* lam_e, float* lam_f)
{
    if (!ctx || nv < 0 || ne < 0 || nf < 0 || nsites < 0 ||
        (nf > 0 && (!face_sites || !face_valid || !sites)) ||
        (nv > 0 && (!ve_off || (ve_off[nv] > 0 && !ve))) ||
        (ne > 0 && (!ef_off || (ef_off[ne] > 0 && !ef))))
        return VC_ERR_INVALID;
    for (int64_t f = 0; f < 2LL * nf; ++f)
        if (face_sites[f] < 0 || face_sites[f] >= nsites)
            return vc_fail(ctx, VC_ERR_INVALID, "face site %d out of range",
                           face_sites[f]);
    const int64_t nve = nv > 0 ? ve_off[nv] : 0, nef = ne > 0 ? ef_off[ne] : 0;
    for (int64_t i = 0; i < nve; ++i)
        if (ve[i] < 0 || ve[i] >= ne)
            return vc_fail(ctx, VC_ERR_INVALID, "vertex-edge entry out of range");
    for (int64_t i = 0; i < nef; ++i)
        if (ef[i] < 0 || ef[i] >= nf)
            return vc_fail(ctx, VC_ERR_INVALID, "edge-face entry out of range");
    for (int32_t v = 0; v < nv; ++v)
        if (ve_off[v] > ve_off[v + 1] || ve_off[v] < 0)
            return vc_fail(ctx, VC_ERR_INVALID, "vertex offsets not ascending");
    for (int32_t e = 0; e < ne; ++e)
        if (ef_off[e] > ef_off[e + 1] || ef_off[e] < 0)
            return vc_fail(ctx, VC_ERR_INVALID, "edge offsets not ascending");
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    Scratch sc;
    float *ds, *dlf, *dle, *dlv;
    int32_t *dfs, *dve, *def;
    uint8_t* dfv;
    int64_t *dveo, *defo;
    int rc = VC_OK;
    if (rc == VC_OK)
        rc = sc.get(ctx, &ds, 3 * (size_t)nsites);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dlf, (size_t)nf);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dle, (size_t)ne);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dlv, (size_t)nv);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dfs, 2 * (size_t)nf);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dfv, (size_t)nf);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dveo, (size_t)nv + 1);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dve, (size_t)nve);
    if (rc == VC_OK)
        rc = sc.get(ctx, &defo, (size_t)ne + 1);
    if (rc == VC_OK)
        rc = sc.get(ctx, &def, (size_t)nef);
    if (rc != VC_OK)
        return rc;
    auto up = [&](void* d, const void* h, size_t bytes) {
        return bytes ? hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, S)
                     : hipSuccess;
    };
    VC_HIP(ctx, up(ds, sites, 12 * (size_t)nsites));
    VC_HIP(ctx, up(dfs, face_sites, 8 * (size_t)nf));
    VC_HIP(ctx, up(dfv, face_valid, (size_t)nf));
    if (nv > 0)
    {
        VC_HIP(ctx, up(dveo, ve_off, 8 * ((size_t)nv + 1)));
        VC_HIP(ctx, up(dve, ve, 4 * (size_t)nve));
    }
    if (ne > 0)
    {
        VC_HIP(ctx, up(defo, ef_off, 8 * ((size_t)ne + 1)));
        VC_HIP(ctx, up(def, ef, 4 * (size_t)nef));
    }
    if (nf > 0)
    {
        face_lambda<<<blocks_for(nf), kBlock, 0, S>>>(ds, dfs, nf, dlf);
        VC_CHECK_LAUNCH(ctx);
    }
    if (ne > 0)
    {
        edge_lambda<<<blocks_for(ne), kBlock, 0, S>>>(dlf, dfv, defo, def, ne,
                                                       dle);
        VC_CHECK_LAUNCH(ctx);
    }
    if (nv > 0 && ne > 0)
    {
        vertex_lambda<<<blocks_for(nv), kBlock, 0, S>>>(dlf, dfv, dveo, dve,
                                                         defo, def, nv, dlv);
        VC_CHECK_LAUNCH(ctx);
    }
    else if (nv > 0)
        VC_HIP(ctx, hipMemsetAsync(dlv, 0, 4 * (size_t)nv, S));
    if (lam_f && nf > 0)
        VC_HIP(ctx, hipMemcpyAsync(lam_f, dlf, 4 * (size_t)nf,
                                   hipMemcpyDeviceToHost, S));
    if (lam_e && ne > 0)
        VC_HIP(ctx, hipMemcpyAsync(lam_e, dle, 4 * (size_t)ne,
                                   hipMemcpyDeviceToHost, S));
    if (lam_v && nv > 0)
        VC_HIP(ctx, hipMemcpyAsync(lam_v, dlv, 4 * (size_t)nv,
                                   hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    return VC_OK;
}

int vc_voro_radii(vc_ctx* ctx, int32_t nv, int32_t nf, const float* vts,
                  const int64_t* face_off, const int32_t* face_vts,
                  const float* sites, int32_t nsites, const int32_t* face_sites,
                  float* radii)
{
    if (!ctx || nv < 0 || nf < 0 || nsites < 0 || (nv > 0 && (!vts || !radii)) ||
        (nf > 0 && (!face_off || !face_vts || !face_sites || !sites)))
        return VC_ERR_INVALID;
    if (nf > 0 && face_off[0] != 0)
        return vc_fail(ctx, VC_ERR_INVALID, "face offsets must start at 0");
    for (int32_t f = 0; f < nf; ++f)
    {
        if (face_off[f + 1] < face_off[f])
            return vc_fail(ctx, VC_ERR_INVALID, "face offsets not ascending");
        if (face_sites[2 * f] < 0 || face_sites[2 * f] >= nsites ||
            face_sites[2 * f + 1] < 0 || face_sites[2 * f + 1] >= nsites)
            return vc_fail(ctx, VC_ERR_INVALID, "face site out of range");
    }
    const int64_t nfv = nf > 0 ? face_off[nf] : 0;
    for (int64_t j = 0; j < nfv; ++j)
        if (face_vts[j] < 0 || face_vts[j] >= nv)
            return vc_fail(ctx, VC_ERR_INVALID, "face vertex out of range");
    if (nv == 0)
        return VC_OK;
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    Scratch sc;
    float *dv, *ds, *dr;
    int32_t *dfv, *dfs, *dlast;
    int64_t* dfo;
    int rc = sc.get(ctx, &dv, 3 * (size_t)nv);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dr, (size_t)nv);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dlast, (size_t)nv);
    if (rc == VC_OK)
        rc = sc.get(ctx, &ds, 3 * (size_t)nsites);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dfs, 2 * (size_t)nf);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dfo, (size_t)nf + 1);
    if (rc == VC_OK)
        rc = sc.get(ctx, &dfv, (size_t)nfv);
    if (rc != VC_OK)
        return rc;
    VC_HIP(ctx, hipMemcpyAsync(dv, vts, 12 * (size_t)nv, hipMemcpyHostToDevice, S));
    VC_HIP(ctx, hipMemsetAsync(dlast, 0xFF, 4 * (size_t)nv, S));
    if (nf > 0)
    {
        if (nsites > 0)
            VC_HIP(ctx, hipMemcpyAsync(ds, sites, 12 * (size_t)nsites,
                                       hipMemcpyHostToDevice, S));
        VC_HIP(ctx, hipMemcpyAsync(dfs, face_sites, 8 * (size_t)nf,
                                   hipMemcpyHostToDevice, S));
        VC_HIP(ctx, hipMemcpyAsync(dfo, face_off, 8 * ((size_t)nf + 1),
                                   hipMemcpyHostToDevice, S));
        if (nfv > 0)
            VC_HIP(ctx, hipMemcpyAsync(dfv, face_vts, 4 * (size_t)nfv,
                                       hipMemcpyHostToDevice, S));
        face_last<<<blocks_for(nf), kBlock, 0, S>>>(dfo, dfv, nf, dlast);
        VC_CHECK_LAUNCH(ctx);
    }
    vertex_radius<<<blocks_for(nv), kBlock, 0, S>>>(dv, ds, dfs, dlast, nv, dr);
    VC_CHECK_LAUNCH(ctx);
    VC_HIP(ctx, hipMemcpyAsync(radii, dr, 4 * (size_t)nv, hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    return VC_OK;
}

int vc_voro_extract_inside_device(
    vc_ctx* ctx, int32_t nv, int32_t ne, int32_t nf, const int32_t* edges,
    const int64_t* face_off, const int32_t* face_vts, const uint8_t* v_valid,
    const float* lam_e, const float* lam_f, int32_t* vts_idx, int64_t vts_cap,
    int64_t* n_vts, int32_t* edges_out, float* em, int64_t edge_cap,
    int64_t* n_edges, int32_t* tris, int32_t* from_fi, float* fm,
    int64_t tri_cap, int64_t* n_tris)
{
    if (!ctx || nv < 0 || ne < 0 || nf < 0 || (ne > 0 && (!edges || !lam_e)) ||
        (nf > 0 && (!face_off || !face_vts || !lam_f)) || (nv > 0 && !v_valid) ||
        !n_vts || !n_edges || !n_tris)
        return VC_ERR_INVALID;
    const int64_t nfv = nf > 0 ? face_off[nf] : 0;
    if ((nf > 0 && face_off[0] != 0) || nfv < 0 || 3 * nfv >= (int64_t)INT32_MAX)
        return VC_ERR_INVALID;
    for (int64_t i = 0; i < 2LL * ne; ++i)
        if (edges[i] < 0 || edges[i] >= nv)
            return VC_ERR_INVALID;
    for (int32_t f = 0; f < nf; ++f)
        if (face_off[f + 1] < face_off[f])
            return VC_ERR_INVALID;
    for (int64_t j = 0; j < nfv; ++j)
        if (face_vts[j] < 0 || face_vts[j] >= nv)
            return VC_ERR_INVALID;
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    Scratch sc;
    int32_t *dE, *dfv;
    int64_t* dfo;
    uint8_t* dval;
    float *dle, *dlf;
    int rc = VC_OK;
    auto get = [&](auto** p, size_t n) {
        if (rc == VC_OK)
            rc = sc.get(ctx, p, n);
    };
    get(&dE, 2 * (size_t)ne + 2);
    get(&dfo, (size_t)nf + 1);
    get(&dfv, (size_t)nfv + 1);
    get(&dval, (size_t)nv + 1);
    get(&dle, (size_t)ne + 1);
    get(&dlf, (size_t)nf + 1);
    if (rc != VC_OK)
        return rc;
    if (ne)
    {
        VC_HIP(ctx, hipMemcpyAsync(dE, edges, 8 * (size_t)ne, hipMemcpyHostToDevice, S));
        VC_HIP(ctx, hipMemcpyAsync(dle, lam_e, 4 * (size_t)ne, hipMemcpyHostToDevice, S));
    }
    if (nf)
    {
        VC_HIP(ctx, hipMemcpyAsync(dfo, face_off, 8 * ((size_t)nf + 1),
                                   hipMemcpyHostToDevice, S));
        VC_HIP(ctx, hipMemcpyAsync(dlf, lam_f, 4 * (size_t)nf, hipMemcpyHostToDevice, S));
        if (nfv)
            VC_HIP(ctx, hipMemcpyAsync(dfv, face_vts, 4 * (size_t)nfv,
                                       hipMemcpyHostToDevice, S));
    }
    if (nv)
        VC_HIP(ctx, hipMemcpyAsync(dval, v_valid, (size_t)nv, hipMemcpyHostToDevice, S));
    vc_voro_xi x;
    struct Free
    {
        vc_voro_xi* x;
        ~Free() { vc_voro_xi_free(x); }
    } fr{&x};
    if ((rc = vc_voro_extract_dev(ctx, nv, ne, nf, dE, dfo, dfv, dval, dle, dlf, &x)) !=
        VC_OK)
        return rc;
    *n_vts = x.nv;
    *n_edges = x.ne;
    *n_tris = x.nt;
    if ((vts_idx && vts_cap < *n_vts) || (edges_out && edge_cap < *n_edges) ||
        (tris && tri_cap < *n_tris))
        return VC_ERR_CAPACITY;
    if (vts_idx && x.nv)
        VC_HIP(ctx, hipMemcpyAsync(vts_idx, x.vi, 4 * (size_t)x.nv, hipMemcpyDeviceToHost, S));
    if (edges_out && x.ne)
        VC_HIP(ctx, hipMemcpyAsync(edges_out, x.eo, 8 * (size_t)x.ne,
                                   hipMemcpyDeviceToHost, S));
    if (em && x.ne)
        VC_HIP(ctx, hipMemcpyAsync(em, x.em, 4 * (size_t)x.ne, hipMemcpyDeviceToHost, S));
    if (tris && x.nt)
        VC_HIP(ctx, hipMemcpyAsync(tris, x.tr, 12 * (size_t)x.nt, hipMemcpyDeviceToHost, S));
    if (from_fi && x.nt)
        VC_HIP(ctx, hipMemcpyAsync(from_fi, x.ff, 4 * (size_t)x.nt, hipMemcpyDeviceToHost, S));
    if (fm && x.nt)
        VC_HIP(ctx, hipMemcpyAsync(fm, x.fm, 4 * (size_t)x.nt, hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    return VC_OK;
}

int vc_cc_finalize_device(vc_ctx* ctx, int32_t nv, int32_t ne_in,
                          const int32_t* edges_in, int32_t nf,
                          const int64_t* face_off, const int32_t* face_vts,
                          int32_t* edges_out, int32_t edge_cap, int32_t* ne_out,
                          int32_t* face_edges_out)
{
    if (!ctx || nv < 0 || ne_in < 0 || nf < 0 || !ne_out ||
        (ne_in && !edges_in) || (nf && (!face_off || !face_vts || !face_edges_out)))
        return VC_ERR_INVALID;
    const int64_t nfv = nf ? face_off[nf] : 0;
    int rc = vc_cc_check_host(ctx, nv, ne_in, edges_in, nf, face_off, face_vts);
    if (rc != VC_OK)
        return rc;
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    Scratch sc;
    const int64_t NC = (int64_t)ne_in + nfv;
    int32_t *dE, *dfv, *deo, *dfe;
    int64_t* dfo;
    auto get = [&](auto** p, size_t n) {
        if (rc == VC_OK)
            rc = sc.get(ctx, p, n);
    };
    get(&dE, 2 * (size_t)ne_in + 2);
    get(&dfo, (size_t)nf + 1);
    get(&dfv, (size_t)nfv + 1);
    get(&deo, 2 * (size_t)NC + 2);
    get(&dfe, (size_t)nfv + 1);
    if (rc != VC_OK)
        return rc;
    if (ne_in)
        VC_HIP(ctx, hipMemcpyAsync(dE, edges_in, 8 * (size_t)ne_in,
                                   hipMemcpyHostToDevice, S));
    if (nf)
    {
        VC_HIP(ctx, hipMemcpyAsync(dfo, face_off, 8 * ((size_t)nf + 1),
                                   hipMemcpyHostToDevice, S));
        if (nfv)
            VC_HIP(ctx, hipMemcpyAsync(dfv, face_vts, 4 * (size_t)nfv,
                                       hipMemcpyHostToDevice, S));
    }
    int32_t ne = 0;
    if ((rc = vc_cc_finalize_dev(ctx, ne_in, dE, nf, dfo, dfv, nfv, deo, dfe, &ne)) !=
        VC_OK)
        return rc;
    *ne_out = ne;
    if ((int64_t)ne > edge_cap || !edges_out)
        return edges_out ? VC_ERR_CAPACITY : VC_OK;
    if (ne)
        VC_HIP(ctx, hipMemcpyAsync(edges_out, deo, 8 * (size_t)ne,
                                   hipMemcpyDeviceToHost, S));
    if (nfv)
        VC_HIP(ctx, hipMemcpyAsync(face_edges_out, dfe, 4 * (size_t)nfv,
                                   hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    return VC_OK;
}

int vc_voro_from_ft(vc_ctx* ctx, vc_vol* vol, float* vts, int64_t vts_cap,
                    int64_t* n_vts, int32_t* faces, int32_t* face_sites,
                    int64_t face_cap, int64_t* n_faces, float* sites,
                    int64_t site_cap, int64_t* n_sites)
{
    if (!ctx || !vol || vol->ctx != ctx || !n_vts || !n_faces || !n_sites ||
        vts_cap < 0 || face_cap < 0 || site_cap < 0)
        return VC_ERR_INVALID;
    if (!vol->edt_done)
        return vc_fail(ctx, VC_ERR_STATE, "vc_voro_from_ft: run vc_edt first");
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    int64_t tot[3];
    int rc;
    {
        Scratch sc;
        uint32_t *tile, *lat, *sid;
        if ((rc = ft_counts(ctx, vol, sc, &tile, &lat, &sid, tot)) != VC_OK)
            return rc;
    }
    *n_faces = tot[0];
    *n_vts = tot[1];
    *n_sites = tot[2];
    if (!vts || !faces || !face_sites || !sites)
        return VC_OK; // size query
    if (tot[1] > vts_cap || tot[0] > face_cap || tot[2] > site_cap)
        return VC_ERR_CAPACITY;
    vc_voro_dev d;
    struct Free
    {
        vc_voro_dev* d;
        ~Free() { vc_voro_dev_free(d); }
    } fr{&d};
    if ((rc = vc_voro_from_ft_dev(ctx, vol, &d)) != VC_OK)
        return rc;
    if (d.nf > 0)
    {
        VC_HIP(ctx, hipMemcpyAsync(faces, d.faces, 16 * (size_t)d.nf,
                                   hipMemcpyDeviceToHost, S));
        VC_HIP(ctx, hipMemcpyAsync(face_sites, d.fs, 8 * (size_t)d.nf,
                                   hipMemcpyDeviceToHost, S));
    }
    if (d.nv > 0)
        VC_HIP(ctx, hipMemcpyAsync(vts, d.vts, 12 * (size_t)d.nv,
                                   hipMemcpyDeviceToHost, S));
    if (d.ns > 0)
        VC_HIP(ctx, hipMemcpyAsync(sites, d.sites, 12 * (size_t)d.ns,
                                   hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    return VC_OK;
}

} // extern "C"

// ---- device-array cores (vc_internal.h)

void vc_voro_dev_free(vc_voro_dev* d)
{
    vc_free(d->vts);
    vc_free(d->faces);
    vc_free(d->fs);
    vc_free(d->sites);
    *d = vc_voro_dev{};
}

int vc_voro_from_ft_dev(vc_ctx* ctx, vc_vol* vol, vc_voro_dev* out)
{
    *out = vc_voro_dev{};
    if (!vol->edt_done)
        return vc_fail(ctx, VC_ERR_STATE, "vc_voro_from_ft: run vc_edt first");
    hipStream_t S = ctx->stream;
    const int64_t N = vol->N;
    const int64_t L = (int64_t)(vol->nx + 1) * (vol->ny + 1) * (vol->nz + 1);
    Scratch sc;
    uint32_t *tile, *lat, *sid;
    int64_t tot[3];
    int rc = ft_counts(ctx, vol, sc, &tile, &lat, &sid, tot);
    if (rc != VC_OK)
        return rc;
    const int64_t F = tot[0], V = tot[1], NS = tot[2];
    if ((rc = own_alloc(ctx, &out->faces, 4 * (size_t)F)) != VC_OK ||
        (rc = own_alloc(ctx, &out->fs, 2 * (size_t)F)) != VC_OK ||
        (rc = own_alloc(ctx, &out->vts, 3 * (size_t)V)) != VC_OK ||
        (rc = own_alloc(ctx, &out->sites, 3 * (size_t)NS)) != VC_OK)
    {
        vc_voro_dev_free(out);
        return rc;
    }
    out->nf = F;
    out->nv = V;
    out->ns = NS;
    const int64_t tiles = (N + kBlock - 1) / kBlock;
    const FtGrid G{vol->occ, vol->ft, vol->nx, vol->ny, vol->nz, vol->pk};
    vft_emit<<<(unsigned)tiles, kBlock, 0, S>>>(G, N, tile, lat, sid, out->faces,
                                                 out->fs);
    vft_vertices<<<blocks_for(L), kBlock, 0, S>>>(vol->nx, vol->ny, vol->nz, lat,
                                                   out->vts);
    vft_sites<<<blocks_for(N), kBlock, 0, S>>>(G, N, sid, out->sites);
    const hipError_t e = hipGetLastError();
    // the scratch (tile / lattice / site flags) is freed on return
    const hipError_t e2 = hipStreamSynchronize(S);
    if (e != hipSuccess || e2 != hipSuccess)
    {
        vc_voro_dev_free(out);
        return vc_fail(ctx, VC_ERR_HIP, "vc_voro_from_ft: %s",
                       hipGetErrorString(e != hipSuccess ? e : e2));
    }
    return VC_OK;
}

int vc_voro_tag_dev(vc_ctx* ctx, vc_vol* vol, const float* pts, int64_t n,
                    uint8_t* valid)
{
    if (n > 0)
        tag_points<<<blocks_for(n), kBlock, 0, ctx->stream>>>(pts, n, vol->occ, vol->nx,
                                                               vol->ny, vol->nz, valid);
    VC_CHECK_LAUNCH(ctx);
    return VC_OK;
}

int vc_voro_radii_dev(vc_ctx* ctx, int32_t nv, int32_t nf, const float* vts,
                      const int64_t* face_off, const int32_t* face_vts,
                      const float* sites, const int32_t* face_sites, int32_t* last,
                      float* r)
{
    hipStream_t S = ctx->stream;
    if (nv == 0)
        return VC_OK;
    VC_HIP(ctx, hipMemsetAsync(last, 0xFF, 4 * (size_t)nv, S));
    if (nf > 0)
        face_last<<<blocks_for(nf), kBlock, 0, S>>>(face_off, face_vts, nf, last);
    vertex_radius<<<blocks_for(nv), kBlock, 0, S>>>(vts, sites, face_sites, last, nv, r);
    VC_CHECK_LAUNCH(ctx);
    return VC_OK;
}

int vc_voro_measures_dev(vc_ctx* ctx, int32_t nv, int32_t ne, int32_t nf,
                         const float* sites, const int32_t* face_sites,
                         const uint8_t* face_valid, const uint32_t* ve_off,
                         const int32_t* ve, const uint32_t* ef_off, const int32_t* ef,
                         float* lam_v, float* lam_e, float* lam_f)
{
    hipStream_t S = ctx->stream;
    if (nf > 0)
        face_lambda<<<blocks_for(nf), kBlock, 0, S>>>(sites, face_sites, nf, lam_f);
    if (ne > 0)
        edge_lambda<<<blocks_for(ne), kBlock, 0, S>>>(lam_f, face_valid, ef_off, ef, ne,
                                                       lam_e);
    if (nv > 0 && ne > 0)
        vertex_lambda<<<blocks_for(nv), kBlock, 0, S>>>(lam_f, face_valid, ve_off, ve,
                                                         ef_off, ef, nv, lam_v);
    else if (nv > 0)
        VC_HIP(ctx, hipMemsetAsync(lam_v, 0, 4 * (size_t)nv, S));
    VC_CHECK_LAUNCH(ctx);
    return VC_OK;
}

void vc_voro_xi_free(vc_voro_xi* x)
{
    vc_free(x->vi);
    vc_free(x->eo);
    vc_free(x->em);
    vc_free(x->tr);
    vc_free(x->ff);
    vc_free(x->fm);
    *x = vc_voro_xi{};
}

int vc_voro_extract_dev(vc_ctx* ctx, int32_t nv, int32_t ne, int32_t nf,
                        const int32_t* dE, const int64_t* dfo, const int32_t* dfv,
                        const uint8_t* dval, const float* dle, const float* dlf,
                        vc_voro_xi* out)
{
    *out = vc_voro_xi{};
    hipStream_t S = ctx->stream;
    Scratch sc;
    uint32_t *vpos, *epos, *toff;
    int rc = VC_OK;
    auto get = [&](auto** p, size_t n) {
        if (rc == VC_OK)
            rc = sc.get(ctx, p, n);
    };
    auto own = [&](auto** p, size_t n) {
        if (rc == VC_OK)
            rc = own_alloc(ctx, p, n);
    };
    get(&vpos, (size_t)nv + 1);
    get(&epos, (size_t)ne + 1);
    get(&toff, (size_t)nf + 1);
    if (rc != VC_OK)
        return rc;
    VC_HIP(ctx, hipMemsetAsync(vpos, 0, 4 * ((size_t)nv + 1), S));
    VC_HIP(ctx, hipMemsetAsync(epos, 0, 4 * ((size_t)ne + 1), S));
    VC_HIP(ctx, hipMemsetAsync(toff, 0, 4 * ((size_t)nf + 1), S));
    const int64_t mx = std::max<int64_t>(std::max<int64_t>(nv, ne), nf);
    if (mx)
        xi_flags<<<blocks_for(mx), kBlock, 0, S>>>(nv, ne, nf, dE, dfo, dfv, dval,
                                                   vpos, epos, toff);
    VC_CHECK_LAUNCH(ctx);
    if ((rc = vc_scan_u32(ctx, vpos, (int64_t)nv + 1)) != VC_OK ||
        (rc = vc_scan_u32(ctx, epos, (int64_t)ne + 1)) != VC_OK ||
        (rc = vc_scan_u32(ctx, toff, (int64_t)nf + 1)) != VC_OK)
        return rc;
    uint32_t cnt[3];
    VC_HIP(ctx, hipMemcpy(&cnt[0], vpos + nv, 4, hipMemcpyDeviceToHost));
    VC_HIP(ctx, hipMemcpy(&cnt[1], epos + ne, 4, hipMemcpyDeviceToHost));
    VC_HIP(ctx, hipMemcpy(&cnt[2], toff + nf, 4, hipMemcpyDeviceToHost));
    const int64_t NV = cnt[0], NE = cnt[1], NT = cnt[2], NC = 3 * NT;
    // table: a power of two >= 2 x the keys
    uint64_t H = 1024;
    while (H < 2 * (uint64_t)(NE + NC) + 1)
        H <<= 1;
    uint64_t* keys;
    unsigned long long* vals;
    uint64_t* ckey;
    uint32_t* keep;
    own(&out->vi, (size_t)NV + 1);
    own(&out->eo, 2 * (size_t)(NE + NC) + 2);
    own(&out->em, (size_t)(NE + NC) + 1);
    own(&out->tr, 3 * (size_t)NT + 3);
    own(&out->ff, (size_t)NT + 1);
    own(&out->fm, (size_t)NT + 1);
    get(&keys, (size_t)H);
    get(&vals, (size_t)H);
    get(&ckey, (size_t)NC + 1);
    get(&keep, (size_t)NC + 1);
    if (rc != VC_OK)
    {
        vc_voro_xi_free(out);
        return rc;
    }
    fill_u64<<<blocks_for((int64_t)H), kBlock, 0, S>>>(keys, (int64_t)H, kEmptyKey);
    fill_u64<<<blocks_for((int64_t)H), kBlock, 0, S>>>(
        reinterpret_cast<uint64_t*>(vals), (int64_t)H, ~0ull);
    const int64_t mve = std::max<int64_t>(nv, ne);
    if (mve)
        xi_emit_ve<<<blocks_for(mve), kBlock, 0, S>>>(nv, ne, dE, dval, vpos, epos, dle,
                                                      out->vi, out->eo, out->em, keys,
                                                      vals, H - 1);
    if (nf)
        xi_tris<<<blocks_for(nf), kBlock, 0, S>>>(nf, dfo, dfv, toff, vpos, dlf, out->tr,
                                                  out->ff, out->fm, ckey, keys, vals,
                                                  H - 1);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemsetAsync(keep, 0, 4 * ((size_t)NC + 1), S);
    if (e == hipSuccess && NC)
    {
        xi_keep<<<blocks_for(NC), kBlock, 0, S>>>(NC, ckey, keys, vals, H - 1, keep);
        e = hipGetLastError();
    }
    if (e != hipSuccess)
    {
        (void)hipStreamSynchronize(S);
        vc_voro_xi_free(out);
        return vc_fail(ctx, VC_ERR_HIP, "inside extraction: %s", hipGetErrorString(e));
    }
    if ((rc = vc_scan_u32(ctx, keep, NC + 1)) != VC_OK)
    {
        vc_voro_xi_free(out);
        return rc;
    }
    uint32_t nk = 0;
    e = hipMemcpy(&nk, keep + NC, 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && NC)
    {
        xi_emit_sides<<<blocks_for(NC), kBlock, 0, S>>>(NC, ckey, keep, vpos, out->ff,
                                                        dlf, NE, out->eo, out->em);
        e = hipGetLastError();
    }
    // the scratch is freed on return
    const hipError_t e2 = hipStreamSynchronize(S);
    if (e != hipSuccess || e2 != hipSuccess)
    {
        vc_voro_xi_free(out);
        return vc_fail(ctx, VC_ERR_HIP, "inside extraction: %s",
                       hipGetErrorString(e != hipSuccess ? e : e2));
    }
    out->nv = NV;
    out->ne = NE + nk;
    out->nt = NT;
    return VC_OK;
}
