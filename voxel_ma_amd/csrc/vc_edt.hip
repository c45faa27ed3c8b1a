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
#include "vc_internal.h"

namespace
{

constexpr int kBlock = 256;
// column tiles are capped at 32 KB of LDS so that several workgroups per CU
// overlap their load and search phases
constexpr size_t kLdsBudget = 32 * 1024;
constexpr uint32_t kOob = 0xFFFFFFFEu; // neighbour outside the grid

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Consecutive tiles (which share cache lines and halos) are given to
// workgroups of the same XCD: block L -> tile (L % 8) * per + L / 8.
__device__ __forceinline__ int64_t xcd_tile(int64_t ntiles)
{
    const int64_t L = blockIdx.x;
    const int64_t per = (ntiles + 7) / 8;
    const int64_t T = (L % 8) * per + L / 8;
    return T < ntiles ? T : -1;
}
// every candidate cost fits in int32: the volume constructor rejects shapes
// whose squared diagonal exceeds INT32_MAX
constexpr int kInf = 0x7FFFFFFF;

__global__ void __launch_bounds__(kBlock)
    edt_pass_z(const uint8_t* __restrict__ occ, int16_t* __restrict__ zq,
               int nz)
{
    extern __shared__ uint64_t bgmask[];
    const int64_t row = (int64_t)blockIdx.x + (int64_t)blockIdx.y * gridDim.x;
    const uint8_t* r = occ + row * nz;
    int16_t* out = zq + row * nz;
    const int nw = (nz + 63) >> 6;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    for (int w = wave; w < nw; w += nwaves)
    {
        int z = (w << 6) + lane;
        bool bg = (z < nz) && r[z] == 0;
        uint64_t m = __ballot(bg);
        if (lane == 0)
            bgmask[w] = m;
    }
    __syncthreads();
    for (int z = threadIdx.x; z < nz; z += blockDim.x)
    {
        const int w = z >> 6, b = z & 63;
        const uint64_t m = bgmask[w];
        int res;
        if ((m >> b) & 1ull)
            res = z;
        else
        {
            int L = -1, R = -1;
            uint64_t lm = m & ((1ull << b) - 1ull); // bits below b
            if (lm)
                L = (w << 6) + 63 - __clzll(lm);
            else
                for (int ww = w - 1; ww >= 0; --ww)
                {
                    uint64_t q = bgmask[ww];
                    if (q)
                    {
                        L = (ww << 6) + 63 - __clzll(q);
                        break;
                    }
                }
            uint64_t rm = m & ~((2ull << b) - 1ull); // bits above b
            if (rm)
                R = (w << 6) + __ffsll((unsigned long long)rm) - 1;
            else
                for (int ww = w + 1; ww < nw; ++ww)
                {
                    uint64_t q = bgmask[ww];
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
        }
        out[z] = (int16_t)res;
    }
}

// columns along y: tile = TZ consecutive z at fixed x, all y
__global__ void __launch_bounds__(kBlock)
    edt_pass_y(const int16_t* __restrict__ zq, uint32_t* __restrict__ yz,
               int nx, int ny, int nz, int TZ)
{
    extern __shared__ int16_t colz[]; // [ny][TZ]
    const int nzt = (nz + TZ - 1) / TZ;
    const int64_t T = xcd_tile((int64_t)nx * nzt);
    if (T < 0)
        return;
    const int z0 = (int)(T % nzt) * TZ;
    const int x = (int)(T / nzt);
    const int tz = threadIdx.x % TZ;
    const int ty0 = threadIdx.x / TZ;
    const int tys = blockDim.x / TZ;
    const int z = z0 + tz;
    const bool zin = z < nz;
    const int64_t base = (int64_t)x * ny * nz;
    // 16 independent loads in flight per thread before the LDS stores
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
            if (yy < ny)
                colz[yy * TZ + tz] = v[u];
        }
    }
    __syncthreads();
    if (!zin)
        return;
    for (int y = ty0; y < ny; y += tys)
    {
        int best = kInf;
        int by = -1;
        {
            int q = colz[y * TZ + tz];
            if (q >= 0)
            {
                best = (z - q) * (z - q);
                by = y;
            }
        }
        for (int d = 1;; ++d)
        {
            const int d2 = d * d;
            if (d2 > best)
                break;
            const bool lo = y - d >= 0, hi = y + d < ny;
            if (!lo && !hi)
                break;
            if (lo)
            {
                int q = colz[(y - d) * TZ + tz];
                if (q >= 0)
                {
                    int v = (z - q) * (z - q) + d2;
                    if (v <= best)
                    {
                        best = v;
                        by = y - d;
                    }
                }
            }
            if (hi)
            {
                int q = colz[(y + d) * TZ + tz];
                if (q >= 0)
                {
                    int v = (z - q) * (z - q) + d2;
                    if (v < best)
                    {
                        best = v;
                        by = y + d;
                    }
                }
            }
        }
        uint32_t outv = VC_NONE;
        if (by >= 0)
            outv = ((uint32_t)by << 16) | (uint32_t)(uint16_t)colz[by * TZ + tz];
        yz[base + (int64_t)y * nz + z] = outv;
    }
}

// columns along x: tile = TZ consecutive z at fixed y, all x
__global__ void __launch_bounds__(kBlock)
    edt_pass_x(const uint32_t* yz, uint32_t* ft,
               int32_t* __restrict__ d2out, int nx, int ny, int nz, int TZ,
               vc_pack pk)
{
    extern __shared__ uint32_t colyz[]; // [nx][TZ]
    const int nzt = (nz + TZ - 1) / TZ;
    const int64_t T = xcd_tile((int64_t)ny * nzt);
    if (T < 0)
        return;
    const int z0 = (int)(T % nzt) * TZ;
    const int y = (int)(T / nzt);
    const int tz = threadIdx.x % TZ;
    const int tx0 = threadIdx.x / TZ;
    const int txs = blockDim.x / TZ;
    const int z = z0 + tz;
    const bool zin = z < nz;
    const int64_t sx = (int64_t)ny * nz;
    const int64_t off = (int64_t)y * nz + z;
    constexpr int U = 16;
    for (int x = tx0; x < nx; x += txs * U)
    {
        uint32_t v[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            const int xx = x + u * txs;
            v[u] = (zin && xx < nx) ? yz[(int64_t)xx * sx + off] : VC_NONE;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            const int xx = x + u * txs;
            if (xx < nx)
                colyz[xx * TZ + tz] = v[u];
        }
    }
    __syncthreads();
    if (!zin)
        return;
    for (int x = tx0; x < nx; x += txs)
    {
        int best = kInf;
        int bxv = -1;
        auto cost = [&](int xx) -> int {
            uint32_t v = colyz[xx * TZ + tz];
            if (v == VC_NONE)
                return kInf;
            int yy = (int)(v >> 16), zz = (int)(v & 0xFFFFu);
            return (y - yy) * (y - yy) + (z - zz) * (z - zz);
        };
        {
            int c = cost(x);
            if (c < kInf)
            {
                best = c;
                bxv = x;
            }
        }
        for (int d = 1;; ++d)
        {
            const int dd = d * d;
            if (dd > best)
                break;
            const bool lo = x - d >= 0, hi = x + d < nx;
            if (!lo && !hi)
                break;
            if (lo)
            {
                int c = cost(x - d);
                if (c < kInf && c + dd <= best)
                {
                    best = c + dd;
                    bxv = x - d;
                }
            }
            if (hi)
            {
                int c = cost(x + d);
                if (c < kInf && c + dd < best)
                {
                    best = c + dd;
                    bxv = x + d;
                }
            }
        }
        const int64_t o = (int64_t)x * sx + off;
        if (bxv < 0)
        {
            ft[o] = VC_NONE;
            d2out[o] = -1;
        }
        else
        {
            uint32_t v = colyz[bxv * TZ + tz];
            ft[o] = vc_pack_xyz(pk, bxv, (int)(v >> 16), (int)(v & 0xFFFFu));
            d2out[o] = (int32_t)best;
        }
    }
}

// Per-voxel measure with an x-sliding window: a 64(z) x 8(y) workgroup walks
// a chunk of kMeasXC slices; FT(x-1), FT(x), FT(x+1) stay in registers and
// the y/z neighbours of slice x come from a 10 x 66 LDS window, so each FT
// value is read about once.  A voxel q is background iff FT(q) == q, so no
// occupancy is read.
constexpr int kMeasZ = 64, kMeasY = 8, kMeasXC = 32;

__global__ void __launch_bounds__(kMeasZ * kMeasY)
    measure_lam2(const uint32_t* __restrict__ ft, uint32_t* __restrict__ lam2,
                 int nx, int ny, int nz, vc_pack pk)
{
    __shared__ uint32_t S[kMeasY + 2][kMeasZ + 2];
    const int nzt = (nz + kMeasZ - 1) / kMeasZ;
    const int nyt = (ny + kMeasY - 1) / kMeasY;
    const int nxc = (nx + kMeasXC - 1) / kMeasXC;
    const int64_t T = xcd_tile((int64_t)nzt * nyt * nxc);
    if (T < 0)
        return;
    const int z0 = (int)(T % nzt) * kMeasZ;
    const int y0 = (int)((T / nzt) % nyt) * kMeasY;
    const int xa = (int)(T / ((int64_t)nzt * nyt)) * kMeasXC;
    const int xb = min(nx, xa + kMeasXC);
    const int tz = threadIdx.x, ty = threadIdx.y;
    const int tid = ty * kMeasZ + tz;
    const int z = z0 + tz, y = y0 + ty;
    const bool valid = z < nz && y < ny;
    const int64_t sx = (int64_t)ny * nz;
    const int64_t row = (int64_t)y * nz + z;
    // halo slots: rows y0-1 and y0+8 (66 each), then columns z0-1 and z0+64
    int hy = 0, hz = 0;
    bool hact = false;
    if (tid < 2 * (kMeasZ + 2))
    {
        hy = tid < kMeasZ + 2 ? -1 : kMeasY;
        hz = (tid % (kMeasZ + 2)) - 1;
        hact = true;
    }
    else if (tid < 2 * (kMeasZ + 2) + 2 * kMeasY)
    {
        const int u = tid - 2 * (kMeasZ + 2);
        hy = u >> 1;
        hz = (u & 1) ? kMeasZ : -1;
        hact = true;
    }
    const int gy = y0 + hy, gz = z0 + hz;
    const bool hin = hact && gy >= 0 && gy < ny && gz >= 0 && gz < nz;
    auto load = [&](int xx) -> uint32_t {
        return (valid && xx >= 0 && xx < nx) ? ft[xx * sx + row] : kOob;
    };
    uint32_t prev = load(xa - 1), cur = load(xa);
    for (int x = xa; x < xb; ++x)
    {
        const uint32_t next = load(x + 1);
        const uint32_t halo =
            hin ? ft[x * sx + (int64_t)gy * nz + gz] : kOob;
        S[ty + 1][tz + 1] = valid ? cur : kOob;
        if (hact)
            S[hy + 1][hz + 1] = halo;
        __syncthreads();
        if (valid)
        {
            uint32_t m = 0;
            const uint32_t me = vc_pack_xyz(pk, x, y, z);
            if (cur != VC_NONE && cur != me)
            {
                int fx, fy, fz;
                vc_unpack_xyz(pk, cur, fx, fy, fz);
                auto visit = [&](uint32_t g, uint32_t q) {
                    if (g == kOob || g == VC_NONE || g == q)
                        return;
                    int gx, gy2, gz2;
                    vc_unpack_xyz(pk, g, gx, gy2, gz2);
                    const uint32_t d = (uint32_t)((fx - gx) * (fx - gx)) +
                                       (uint32_t)((fy - gy2) * (fy - gy2)) +
                                       (uint32_t)((fz - gz2) * (fz - gz2));
                    m = d > m ? d : m;
                };
                visit(prev, vc_pack_xyz(pk, x - 1, y, z));
                visit(next, vc_pack_xyz(pk, x + 1, y, z));
                visit(S[ty][tz + 1], vc_pack_xyz(pk, x, y - 1, z));
                visit(S[ty + 2][tz + 1], vc_pack_xyz(pk, x, y + 1, z));
                visit(S[ty + 1][tz], vc_pack_xyz(pk, x, y, z - 1));
                visit(S[ty + 1][tz + 2], vc_pack_xyz(pk, x, y, z + 1));
            }
            lam2[x * sx + row] = m;
        }
        __syncthreads();
        prev = cur;
        cur = next;
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

int pick_tile(int n, size_t elem)
{
    int tz = 64;
    while (tz > 1 && (size_t)tz * (size_t)n * elem > kLdsBudget)
        tz >>= 1;
    return tz;
}

} // namespace

int vc_edt_run(vc_ctx* ctx, vc_vol* v)
{
    VC_HIP(ctx, hipSetDevice(ctx->device));
    const size_t N = (size_t)v->N;
    int rc;
    if (!v->edt && (rc = vc_alloc(ctx, (void**)&v->edt, 4 * N)) != VC_OK)
        return rc;
    if (!v->ft && (rc = vc_alloc(ctx, (void**)&v->ft, 4 * N)) != VC_OK)
        return rc;
    // one full column per workgroup at least (TZ = 1 may use up to 64 KB)
    if ((size_t)v->nx * 4 > 65536 || (size_t)v->ny * 2 > 65536)
        return vc_fail(ctx, VC_ERR_INVALID, "axis too long for the LDS tiles");
    // No temporaries: pass z writes int16 z' into the first half of the edt
    // buffer, pass y writes packed (y', z') into the ft buffer, and pass x
    // rewrites ft in place (each workgroup stages exactly the elements it
    // later overwrites) and overwrites edt once every z' has been consumed.
    int16_t* zq = reinterpret_cast<int16_t*>(v->edt);
    uint32_t* yz = v->ft;
    dim3 gz((unsigned)v->ny, (unsigned)v->nx);
    size_t lz = sizeof(uint64_t) * (size_t)((v->nz + 63) / 64);
    edt_pass_z<<<gz, kBlock, lz, ctx->stream>>>(v->occ, zq, v->nz);
    VC_CHECK_LAUNCH(ctx);
    {
        int tz = pick_tile(v->ny, 2);
        const int64_t tiles = (int64_t)v->nx * ((v->nz + tz - 1) / tz);
        const unsigned g = (unsigned)(8 * ((tiles + 7) / 8));
        edt_pass_y<<<g, kBlock, (size_t)tz * v->ny * 2, ctx->stream>>>(
            zq, yz, v->nx, v->ny, v->nz, tz);
        VC_CHECK_LAUNCH(ctx);
    }
    {
        int tz = pick_tile(v->nx, 4);
        const int64_t tiles = (int64_t)v->ny * ((v->nz + tz - 1) / tz);
        const unsigned g = (unsigned)(8 * ((tiles + 7) / 8));
        edt_pass_x<<<g, kBlock, (size_t)tz * v->nx * 4, ctx->stream>>>(
            yz, v->ft, v->edt, v->nx, v->ny, v->nz, tz, v->pk);
        VC_CHECK_LAUNCH(ctx);
    }
    v->edt_done = true;
    return VC_OK;
}

int vc_measure_run(vc_ctx* ctx, vc_vol* v)
{
    VC_HIP(ctx, hipSetDevice(ctx->device));
    int rc;
    if (!v->lam2 &&
        (rc = vc_alloc(ctx, (void**)&v->lam2, 4 * (size_t)v->N)) != VC_OK)
        return rc;
    const int64_t tiles = (int64_t)((v->nz + kMeasZ - 1) / kMeasZ) *
                          ((v->ny + kMeasY - 1) / kMeasY) *
                          ((v->nx + kMeasXC - 1) / kMeasXC);
    const unsigned g = (unsigned)(8 * ((tiles + 7) / 8));
    measure_lam2<<<g, dim3(kMeasZ, kMeasY), 0, ctx->stream>>>(
        v->ft, v->lam2, v->nx, v->ny, v->nz, v->pk);
    VC_CHECK_LAUNCH(ctx);
    v->measure_kind = 1;
    return VC_OK;
}

int vc_ft_to_linear(vc_ctx* ctx, vc_vol* v, int32_t* dst)
{
    unsigned blocks = (unsigned)((v->N + kBlock - 1) / kBlock);
    ft_linear<<<blocks, kBlock, 0, ctx->stream>>>(v->ft, dst, v->ny, v->nz,
                                                  v->N, v->pk);
    VC_CHECK_LAUNCH(ctx);
    return VC_OK;
}
