// vc_unet.hip -- the surrogate network of include/vc_surrogate.h, gfx950.
//
// Tensors are channels last, [X][Y][Z][C] with C a multiple of 8, BF16.
// A 3^3 convolution is an implicit GEMM: M = voxels, N = output channels,
// K = (tap, 8-channel chunk) pairs.  One v_mfma_f32_16x16x32_bf16 consumes
// four pairs (lane group g = lane / 16 holds one), so any channel count that
// is a multiple of 8 tiles K exactly; the input is staged in LDS per output
// brick (conv_fwd_lds).  The input-gradient pass is the same kernel with the
// taps flipped and the channel roles swapped; the weight gradient stages 64-voxel
// blocks of the layer input and output gradient in LDS and reduces over
// voxels on MFMA into per-chunk partials, summed by a second kernel (fixed
// order: the gradients are deterministic).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "vc_internal.h"
#include "../../include/vc_surrogate.h"

namespace
{

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kBlock = 256;
constexpr float kScale = 32.0f;
constexpr float kEpsT = 1.0f / 32.0f;
constexpr float kErrW = 0.1f;
constexpr float kErrFloor = 1e-4f;
constexpr int kHalo = VC_SURROGATE_HALO;
#ifndef VC_CONV_BX
#define VC_CONV_BX 4
#endif
#ifndef VC_CONV_BY
#define VC_CONV_BY 4
#endif
#ifndef VC_CONV_NZT
#define VC_CONV_NZT 1
#endif
#ifndef VC_CONV_SC
#define VC_CONV_SC 2
#endif
#ifndef VC_CONV_COARSE
#define VC_CONV_COARSE 1
#endif
#ifndef VC_CONV_FUSE01
#define VC_CONV_FUSE01 1
#endif
#ifndef VC_CONV_FUSED_HEAD_CORE
#define VC_CONV_FUSED_HEAD_CORE 1
#endif
#ifndef VC_CONV_GX
#define VC_CONV_GX 4
#endif
#ifndef VC_CONV_UNROLL
#define VC_CONV_UNROLL 4
#endif


struct LayerSpec
{
    int cin, cout, level;
};
// L0a L0b L1a L1b L2a L2b L3a L3b L4a L4b L5a L5b L6a L6b
constexpr LayerSpec kLayers[14] = {
    {2, 16, 0},   {16, 16, 0}, {16, 32, 1}, {32, 32, 1}, {32, 48, 2},
    {48, 48, 2},  {48, 64, 3}, {64, 64, 3}, {112, 48, 2}, {48, 48, 2},
    {80, 32, 1},  {32, 32, 1}, {48, 16, 0}, {16, 16, 0}};
constexpr int kHeadIn = 16, kHeadOut = 2;

constexpr int cinp_of(int cin) { return (cin + 7) / 8 * 8; }
constexpr int round4(int x) { return (x + 3) / 4 * 4; }

int64_t param_offset(int layer) // offset of W of a layer (14 = head)
{
    int64_t o = 0;
    for (int l = 0; l < layer; ++l)
        o += 27LL * kLayers[l].cin * kLayers[l].cout + kLayers[l].cout;
    return o;
}
int64_t total_params()
{
    return param_offset(14) + kHeadIn * kHeadOut + kHeadOut;
}

uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// ------------------------------------------------------------ convolution
// out[v][co] = act(sum_{tap, ci} in[v + off(tap)][ci] W[tap][ci][co] + b[co])
// W: [kc][Cout] of 8-channel chunks (chunk k = tap * cin8 + c8).
// Region voxel (x, y, z) is padded-volume voxel (o + x, ...); outputs
// outside the padded volume [0, P) are written as 0, so that a tile region
// reproduces the zero padding of the whole-volume pass at every layer.
struct Box
{
    int ox, oy, oz, px, py, pz;
};

// The convolution, with the input staged in LDS.  A workgroup (4
// wavefronts) owns an output brick of BX x BY rows (x, y) by 16 NZT voxels
// (z), i.e. BX BY NZT tiles of 16 voxels; wavefront w owns tiles
// w MT .. w MT + MT - 1 (tile t: row t / NZT, z segment t % NZT).  The
// brick's input region (one voxel of halo) is staged SC chunks of 8
// channels at a time, chunk-major (16 B per voxel and chunk, so the 16 lanes
// of a tile read consecutive entries).  Within a slice the K pairs (tap,
// chunk) run tap-major, four per MFMA (lane group g takes pair 4 j + g), so a
// layer with fewer than 4 chunks still fills K; the voxel operand comes
// from LDS, the weight operand (that tap and chunk, shared by every
// workgroup) from global memory, each reused by the wavefront's MT tiles.
//
// kCat: the input is the decoder's concatenation [up | skip] without the
// concatenated tensor: channels below c1 are the coarse level's (nearest
// upsampled: voxel (x, y, z) reads (x/2, y/2, z/2) of `in`, c1 channels),
// the rest the skip tensor `in2` (cinp - c1 channels).
// The MFMAs of one staged slice of S input chunks: K pairs p = tap * S + c,
// four per MFMA (lane group g takes pair p0 + g).  S is a compile-time
// constant here, so a pair's tap and chunk, and the tap's offsets, are
// constant-divisor arithmetic (the run-time division by S had cost ~30 VALU
// instructions per MFMA).
//
// The weights are the MFMA's first operand and the voxels its second, so a
// lane ends with 4 consecutive output channels of one voxel (channel
// n 16 + 4 g + r of voxel i of the tile) and writes them with one 8-byte store.
//
// With two chunks per slice (every layer's slices: all channel counts here
// are multiples of 16), lane group g takes tap 2 j + (g >> 1) and chunk g & 1
// of MFMA j; tap 2 j's region offset is a constant after unrolling, and tap
// 2 j + 1 lies one of three fixed steps further (+1 in z, or the wrap into
// the next y or x row), so a lane's LDS address is one of three registers
// plus an immediate, and the weights of MFMA j sit at a wave-uniform offset.
constexpr int conv_tap_off(int tap, int RY, int RZ)
{
    return ((tap / 9 - 1) * RY + (tap / 3) % 3 - 1) * RZ + tap % 3 - 1;
}

template <int NT, int MT, int NR, int RY, int RZ, int S>
__device__ __forceinline__ void conv_slice(const bf16x8* __restrict__ reg,
                                           const int (&tb)[MT],
                                           const bf16x8* __restrict__ W, int s0,
                                           int cin8, f32x4 (&acc)[MT][NT])
{
    constexpr int Cout = NT * 16;
    constexpr int NP = 27 * S;
    const int lane = threadIdx.x & 63, i = lane & 15, g = lane >> 4;
    const bf16x8 zero8 = {};
    if constexpr (S == 2)
    {
        constexpr int NG = (NP + 3) / 4;
        const int h = g >> 1, c = g & 1;
        const int lb = c * NR + i;
        const int st[3] = {h, h * (RZ - 2), h * (RY * RZ - 2 * RZ - 2)};
        const int wl = (h * cin8 + c) * Cout + i;
#pragma unroll
        for (int j = 0; j < NG; ++j)
        {
            const int t0 = 2 * j;
            const bool v = 4 * j + g < NP; // false only for tap 27 of the last MFMA
            const int step = t0 % 3 < 2 ? 0 : ((t0 / 3) % 3 < 2 ? 1 : 2);
            const int ro = lb + conv_tap_off(t0, RY, RZ) + (v ? st[step] : 0);
            const bf16x8* Wj = W + (t0 * cin8 + s0) * Cout;
            bf16x8 a[MT];
#pragma unroll
            for (int m = 0; m < MT; ++m)
                a[m] = v ? reg[ro + tb[m]] : zero8;
#pragma unroll
            for (int n = 0; n < NT; ++n)
            {
                const bf16x8 b = v ? Wj[wl + n * 16] : zero8;
#pragma unroll
                for (int m = 0; m < MT; ++m)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b, a[m], acc[m][n],
                                                                        0, 0, 0);
            }
        }
    }
    else
    {
#pragma unroll VC_CONV_UNROLL
        for (int p0 = 0; p0 < NP; p0 += 4)
        {
            const int p = p0 + g;
            const bool v = p < NP;
            const int tap = v ? p / S : 13, c = v ? p % S : 0;
            const int ro = c * NR + conv_tap_off(tap, RY, RZ) + i;
            const int wo = (tap * cin8 + s0 + c) * Cout + i;
            bf16x8 a[MT];
#pragma unroll
            for (int m = 0; m < MT; ++m)
                a[m] = v ? reg[ro + tb[m]] : zero8;
#pragma unroll
            for (int n = 0; n < NT; ++n)
            {
                const bf16x8 b = v ? W[wo + n * 16] : zero8;
#pragma unroll
                for (int m = 0; m < MT; ++m)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b, a[m], acc[m][n],
                                                                        0, 0, 0);
            }
        }
    }
}

// The decoder concatenation's coarse-level chunks (channels below c1: the
// nearest up-sampling of the level below) are staged at the coarse level: the
// fine region voxel (rx, ry, rz) (fine coordinate x0 - 1 + rx, ...) is the
// coarse voxel ((x0 - 1 + rx) >> 1, ...), i.e. coarse region voxel ((rx + 1) >> 1,
// (ry + 1) >> 1, (rz + 1) >> 1) for an even brick origin, so a 4 x 4 x 16 brick
// stages 4 x 4 x 10 coarse voxels per chunk instead of 6 x 6 x 18 fine ones
// (the same values: each coarse voxel once instead of up to 8 times).
template <int BX, int BY, int BZ>
struct CoarseRegion
{
    static constexpr int RX = BX / 2 + 2, RY = BY / 2 + 2, RZ = BZ / 2 + 2,
                         NR = RX * RY * RZ;
};

template <int BX, int BY, int BZ, int SC>
__device__ __forceinline__ void conv_stage_coarse(bf16x8* __restrict__ reg,
                                                  const bf16* __restrict__ in, int c1, int s0,
                                                  int x0, int y0, int z0, int X, int Y, int Z)
{
    using CR = CoarseRegion<BX, BY, BZ>;
    constexpr int NE = SC * CR::NR, NIT = (NE + kBlock - 1) / kBlock, RPI = kBlock / SC;
    static_assert(kBlock % SC == 0, "chunk per thread");
    const int c = threadIdx.x % SC, r0 = threadIdx.x / SC;
    const int Xc = X >> 1, Yc = Y >> 1, Zc = Z >> 1;
    const int cx0 = (x0 - 1) >> 1, cy0 = (y0 - 1) >> 1, cz0 = (z0 - 1) >> 1;
    bf16x8 v[NIT];
    uint32_t okm = 0;
#pragma unroll
    for (int k = 0; k < NIT; ++k)
    {
        const int r = min(r0 + k * RPI, CR::NR - 1);
        const int cz = cz0 + r % CR::RZ, cy = cy0 + (r / CR::RZ) % CR::RY,
                  cx = cx0 + r / (CR::RZ * CR::RY);
        const bool ok = (unsigned)cx < (unsigned)Xc && (unsigned)cy < (unsigned)Yc &&
                        (unsigned)cz < (unsigned)Zc;
        okm |= ok ? 1u << k : 0u;
        const int64_t o = ok ? (((int64_t)cx * Yc + cy) * Zc + cz) * c1 + (s0 + c) * 8 : 0;
        v[k] = *reinterpret_cast<const bf16x8*>(in + o);
    }
    const bf16x8 zero8 = {};
#pragma unroll
    for (int k = 0; k < NIT; ++k)
    {
        const int r = r0 + k * RPI;
        if (r < CR::NR)
            reg[c * CR::NR + r] = (okm >> k) & 1u ? v[k] : zero8;
    }
}

// The MFMAs of such a slice (two chunks; the pairing of conv_slice's two-chunk
// path, the voxel operands read through the coarse addressing).  Wavefront w
// holds brick x row w, tile m brick y row m.
template <int NT, int MT, int BX, int BY, int BZ>
__device__ __forceinline__ void conv_slice_coarse(const bf16x8* __restrict__ reg,
                                                  const bf16x8* __restrict__ W, int s0,
                                                  int cin8, f32x4 (&acc)[MT][NT])
{
    using CR = CoarseRegion<BX, BY, BZ>;
    constexpr int Cout = NT * 16, NP = 27 * 2, NG = (NP + 3) / 4;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = lane & 15, g = lane >> 4, h = g >> 1, c = g & 1;
    const bf16x8 zero8 = {};
    // coarse z of fine region z 1 + i + dz, dz = -1, 0, 1
    const int zl[3] = {(i + 1) >> 1, (i + 2) >> 1, (i + 3) >> 1};
    const int lc = c * CR::NR;
    const int wl = (h * cin8 + c) * Cout + i;
#pragma unroll
    for (int j = 0; j < NG; ++j)
    {
        const int t0 = 2 * j, t1 = 2 * j + 1 < 27 ? 2 * j + 1 : 2 * j;
        const bool v = 4 * j + g < NP; // false only for tap 27 of the last MFMA
        const int dx0 = t0 / 9 - 1, dy0 = (t0 / 3) % 3 - 1, dz0 = t0 % 3 - 1;
        const int dx1 = t1 / 9 - 1, dy1 = (t1 / 3) % 3 - 1, dz1 = t1 % 3 - 1;
        const int x0r = ((wv + 2 + dx0) >> 1) * (CR::RY * CR::RZ) + zl[dz0 + 1];
        const int x1r = ((wv + 2 + dx1) >> 1) * (CR::RY * CR::RZ) + zl[dz1 + 1];
        bf16x8 a[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m)
        {
            const int a0 = x0r + ((m + 2 + dy0) >> 1) * CR::RZ;
            const int a1 = x1r + ((m + 2 + dy1) >> 1) * CR::RZ;
            a[m] = v ? reg[lc + (h ? a1 : a0)] : zero8;
        }
        const bf16x8* Wj = W + (t0 * cin8 + s0) * Cout;
#pragma unroll
        for (int n = 0; n < NT; ++n)
        {
            const bf16x8 b = v ? Wj[wl + n * 16] : zero8;
#pragma unroll
            for (int m = 0; m < MT; ++m)
                acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b, a[m], acc[m][n], 0, 0,
                                                                    0);
        }
    }
}

// Bias, ReLU and the zeroing outside the padded volume for one wavefront's
// MT tiles: lane (i, g) holds channels n 16 + 4 g .. + 3 of voxel i.  With
// off32 the output offsets from the brick's first voxel fit 31 bits.
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
template <int NT, int MT, int BY, int NZT, bool kRelu>
__device__ __forceinline__ void conv_store(const f32x4 (&acc)[MT][NT], bf16* __restrict__ out,
                                           const float* __restrict__ bias, int x0, int y0,
                                           int z0, int X, int Y, int Z, const Box& box,
                                           int off32)
{
    constexpr int Cout = NT * 16;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = lane & 15, g = lane >> 4;
    f32x4 bv[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            bv[n][r] = bias ? bias[n * 16 + 4 * g + r] : 0.0f;
    bf16* ob = out + (((int64_t)x0 * Y + y0) * Z + z0) * Cout + 4 * g;
#pragma unroll
    for (int m = 0; m < MT; ++m)
    {
        const int t = wv * MT + m, row = t / NZT, seg = t % NZT;
        const int x = x0 + row / BY, y = y0 + row % BY, z = z0 + 16 * seg + i;
        if (x >= X || y >= Y || z >= Z)
            continue;
        const bool inbox = (unsigned)(box.ox + x) < (unsigned)box.px &&
                           (unsigned)(box.oy + y) < (unsigned)box.py &&
                           (unsigned)(box.oz + z) < (unsigned)box.pz;
        const int lo = ((row / BY) * Y + row % BY) * Z + 16 * seg + i;
        bf16* o = off32 ? ob + lo * Cout : ob + (int64_t)lo * Cout;
#pragma unroll
        for (int n = 0; n < NT; ++n)
        {
            f32x4 val = acc[m][n] + bv[n];
            if (kRelu)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    val[r] = fmaxf(val[r], 0.0f);
            u32x2 q = __builtin_bit_cast(u32x2, __builtin_convertvector(val, bf16x4));
            q = inbox ? q : u32x2{0u, 0u};
            *reinterpret_cast<u32x2*>(o + n * 16) = q;
        }
    }
}

// The brick of a workgroup.  Workgroups are dealt round-robin to the 8 XCDs
// (each with its own L2); XCD k takes the k-th eighth of the bricks in order
// (z fastest), so bricks sharing halo planes of their input regions share an
// L2.  The grid is 8 ceil(nb / 8) workgroups; -1 past the last brick.
__device__ __forceinline__ int conv_brick(int nb)
{
    const int per = (nb + 7) >> 3;
    const int b = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    return b < nb ? b : -1;
}
inline unsigned conv_grid(int64_t nb)
{
    return (unsigned)(8 * ((nb + 7) / 8));
}

// Staging of one slice (SC chunks) of a brick's input region.  Thread t
// stages chunk t % SC of region voxels t / SC + k kBlock / SC.  A brick whose
// region lies inside the volume (every brick but those on its faces) takes
// the first branch: no bounds checks and 32-bit offsets from a wave-uniform
// base (the host passes off32 when a region's extent fits them), with all of
// a thread's loads issued before its LDS stores.
// The last layer of inference with the head fused: its activations are not
// stored; lane group g takes tile g, gathers the 16 BF16-rounded channels of
// voxel i (the values the stored layer would hold) and evaluates the head as
// head_fwd does (out = bias, then one FMA per channel in channel order), so
// the outputs are those of the unfused pair.
// With `core` (an inference's lam / rel arrays), the outputs of the region's
// core voxels that lie in the volume are also written there as copy_core
// writes them (lam = out0 * 32, rel = exp(out1)).
struct CoreOut
{
    float* lam;
    float* rel;
    int nx, ny, nz, cx0, cy0, cz0, cx1, cy1, cz1;
};
template <int MT, int BY, int NZT>
__device__ __forceinline__ void conv_store_head(const f32x4 (&acc)[MT][1],
                                                const float* __restrict__ bias,
                                                const float* __restrict__ Wh,
                                                float* __restrict__ hout, int x0, int y0,
                                                int z0, int X, int Y, int Z, const Box& box,
                                                const CoreOut& core)
{
    static_assert(MT == 4, "one tile per lane group");
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = lane & 15, g = lane >> 4;
    float bv[4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
        bv[r] = bias[4 * g + r];
    float H[16];
#pragma unroll
    for (int m = 0; m < MT; ++m)
    {
        const int t = wv * MT + m, row = t / NZT, seg = t % NZT;
        const int x = x0 + row / BY, y = y0 + row % BY, z = z0 + 16 * seg + i;
        const bool inbox = (unsigned)(box.ox + x) < (unsigned)box.px &&
                           (unsigned)(box.oy + y) < (unsigned)box.py &&
                           (unsigned)(box.oz + z) < (unsigned)box.pz;
        float h[4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
            h[r] = inbox ? (float)(bf16)fmaxf(acc[m][0][r] + bv[r], 0.0f) : 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int r = 0; r < 4; ++r)
            {
                const float v = __shfl(h[r], i + 16 * k);
                if (g == m)
                    H[4 * k + r] = v;
            }
    }
    const int t = wv * MT + g, row = t / NZT, seg = t % NZT;
    const int x = x0 + row / BY, y = y0 + row % BY, z = z0 + 16 * seg + i;
    if (x >= X || y >= Y || z >= Z)
        return;
    float o0 = Wh[kHeadIn * kHeadOut], o1 = Wh[kHeadIn * kHeadOut + 1];
#pragma unroll
    for (int c = 0; c < kHeadIn; ++c)
    {
        o0 = __builtin_fmaf(H[c], Wh[c * 2], o0);
        o1 = __builtin_fmaf(H[c], Wh[c * 2 + 1], o1);
    }
    const int64_t v = ((int64_t)x * Y + y) * Z + z;
    *reinterpret_cast<float2*>(hout + 2 * v) = make_float2(o0, o1);
    const int gx = box.ox + x, gy = box.oy + y, gz = box.oz + z;
    if (core.lam && gx >= core.cx0 && gx < core.cx1 && gy >= core.cy0 && gy < core.cy1 &&
        gz >= core.cz0 && gz < core.cz1 && gx < core.nx && gy < core.ny && gz < core.nz)
    {
        const int64_t i = ((int64_t)gx * core.ny + gy) * core.nz + gz;
        core.lam[i] = o0 * kScale;
        core.rel[i] = expf(o1);
    }
}

template <int NR, int RY, int RZ, int SC, bool kCat>
__device__ __forceinline__ void conv_stage(bf16x8* __restrict__ reg, const bf16* __restrict__ in,
                                           const bf16* __restrict__ in2, int c1, int cinp,
                                           int s0, int x0, int y0, int z0, int X, int Y,
                                           int Z, bool inner)
{
    constexpr int NIT = (NR * SC + kBlock - 1) / kBlock, RPI = kBlock / SC;
    static_assert(kBlock % SC == 0, "chunk per thread");
    const int c = threadIdx.x % SC, r0 = threadIdx.x / SC;
    const int ch = s0 + c, c18 = c1 >> 3, c2 = cinp - c1;
    const bf16x8 zero8 = {};
    if (inner)
    {
        const bf16* base;
        int cs, ys, zs, ax = 0, ay = 0, az = 0;
        if (kCat && ch < c18)
        {
            // coarse voxel (x >> 1, ...) of the region voxel: offsets from
            // the coarse voxel of the region origin
            ax = (x0 - 1) & 1;
            ay = (y0 - 1) & 1;
            az = (z0 - 1) & 1;
            ys = Y >> 1;
            zs = Z >> 1;
            cs = c1;
            base = in + ((((int64_t)((x0 - 1) >> 1)) * ys + ((y0 - 1) >> 1)) * zs +
                         ((z0 - 1) >> 1)) * c1 + ch * 8;
        }
        else
        {
            ys = Y;
            zs = Z;
            cs = kCat ? c2 : cinp;
            base = (kCat ? in2 : in) + ((((int64_t)(x0 - 1)) * Y + (y0 - 1)) * Z + (z0 - 1)) * cs +
                   (kCat ? ch - c18 : ch) * 8;
        }
        const int sh = kCat && ch < c18 ? 1 : 0;
        bf16x8 v[NIT];
#pragma unroll
        for (int k = 0; k < NIT; ++k)
        {
            const int r = min(r0 + k * RPI, NR - 1);
            const int rz = r % RZ, rxy = r / RZ, rx = rxy / RY, ry = rxy % RY;
            const int off = ((((ax + rx) >> sh) * ys + ((ay + ry) >> sh)) * zs +
                             ((az + rz) >> sh)) * cs;
            v[k] = *reinterpret_cast<const bf16x8*>(base + off);
        }
#pragma unroll
        for (int k = 0; k < NIT; ++k)
        {
            const int r = r0 + k * RPI;
            if (k < NIT - 1 || r < NR)
                reg[c * NR + r] = v[k];
        }
        return;
    }
    for (int r = r0; r < NR; r += RPI)
    {
        const int rz = r % RZ, rxy = r / RZ;
        const int xx = x0 - 1 + rxy / RY, yy = y0 - 1 + rxy % RY, zz = z0 - 1 + rz;
        bf16x8 v = zero8;
        if ((unsigned)xx < (unsigned)X && (unsigned)yy < (unsigned)Y &&
            (unsigned)zz < (unsigned)Z)
        {
            if (kCat && ch < c18)
                v = *reinterpret_cast<const bf16x8*>(
                    in + ((((int64_t)(xx >> 1) * (Y >> 1)) + (yy >> 1)) * (Z >> 1) +
                          (zz >> 1)) * c1 + ch * 8);
            else if (kCat)
                v = *reinterpret_cast<const bf16x8*>(
                    in2 + (((int64_t)xx * Y + yy) * Z + zz) * c2 + (ch - c18) * 8);
            else
                v = *reinterpret_cast<const bf16x8*>(
                    in + (((int64_t)xx * Y + yy) * Z + zz) * cinp + ch * 8);
        }
        reg[c * NR + r] = v;
    }
}

template <int NT, int BX, int BY, int NZT, int SC, bool kCat, bool kRelu, bool kHead = false>
__global__ void __launch_bounds__(kBlock)
    conv_fwd_lds(const bf16* __restrict__ in, const bf16* __restrict__ in2, int c1,
                 int cinp, bf16* __restrict__ out, const bf16x8* __restrict__ W,
                 const float* __restrict__ bias, int X, int Y, int Z, Box box, int off32,
                 const float* __restrict__ Wh = nullptr, float* __restrict__ hout = nullptr,
                 CoreOut core = CoreOut{})
{
    constexpr int Cout = NT * 16;
    constexpr int MT = BX * BY * NZT / 4;
    static_assert(MT * 4 == BX * BY * NZT, "tiles per wavefront");
    constexpr int BZ = 16 * NZT, RY = BY + 2, RZ = BZ + 2;
    constexpr int NR = (BX + 2) * RY * RZ;
    __shared__ bf16x8 reg[SC * NR];
    const int wv = threadIdx.x >> 6;
    const int cin8 = cinp >> 3;
    const int nbz = (Z + BZ - 1) / BZ, nby = (Y + BY - 1) / BY;
    const int nbx = (X + BX - 1) / BX;
    int blk = conv_brick(nbz * nby * nbx);
    if (blk < 0)
        return;
    const int bz = blk % nbz;
    blk /= nbz;
    // x groups of VC_CONV_GX bricks, y within a group, then the groups
    const int q = blk / (nby * VC_CONV_GX), rem = blk % (nby * VC_CONV_GX);
    const int gsz = min(VC_CONV_GX, nbx - q * VC_CONV_GX);
    const int by = rem / gsz, bx = q * VC_CONV_GX + rem % gsz;
    const int x0 = BX * bx, y0 = BY * by, z0 = bz * BZ;
    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n)
            acc[m][n] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // region offset of each tile's first voxel (row + 1, z segment + 1)
    int tb[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
    {
        const int t = wv * MT + m, row = t / NZT, seg = t % NZT;
        tb[m] = ((row / BY + 1) * RY + (row % BY + 1)) * RZ + 1 + 16 * seg;
    }
    // the chunk count is a multiple of SC (checked by the host)
    const bool inner = off32 && x0 >= 1 && y0 >= 1 && z0 >= 1 && x0 + BX < X &&
                       y0 + BY < Y && z0 + BZ < Z;
    constexpr bool kCoarse = kCat && SC == 2 && NZT == 1 && MT == BY && BX % 2 == 0 &&
                             BY % 2 == 0 && VC_CONV_COARSE;
    int s0 = 0;
    if constexpr (kCoarse)
        for (; s0 + SC <= (c1 >> 3); s0 += SC)
        {
            __syncthreads(); // the previous slice's reads are done
            conv_stage_coarse<BX, BY, BZ, SC>(reg, in, c1, s0, x0, y0, z0, X, Y, Z);
            __syncthreads();
            conv_slice_coarse<NT, MT, BX, BY, BZ>(reg, W, s0, cin8, acc);
        }
    for (; s0 < cin8; s0 += SC)
    {
        __syncthreads(); // the previous slice's reads are done
        conv_stage<NR, RY, RZ, SC, kCat>(reg, in, in2, c1, cinp, s0, x0, y0, z0, X, Y, Z,
                                         inner);
        __syncthreads();
        conv_slice<NT, MT, NR, RY, RZ, SC>(reg, tb, W, s0, cin8, acc);
    }
    if constexpr (kHead)
        conv_store_head<MT, BY, NZT>(acc, bias, Wh, hout, x0, y0, z0, X, Y, Z, box, core);
    else
        conv_store<NT, MT, BY, NZT, kRelu>(acc, out, bias, x0, y0, z0, X, Y, Z, box, off32);
}

// Weight gradient partials.  Block (chunk, tg): taps (dx, dy) = (tg/3 - 1,
// tg%3 - 1), dz = -1, 0, 1; voxels [chunk * vbpc * 64, ...) in 64-voxel
// blocks.  part[chunk][tap][ci][co] = sum_v x[v + off(tap)][ci] dzt[v][co].
// MFMA rows = output channels (A = dzt^T), columns = (dz, 16 input
// channels) (B = shifted x), K = 32 voxels.
template <int RT, int CT>
__global__ void __launch_bounds__(kBlock)
    conv_wgrad(const bf16* __restrict__ x, int cinp,
               const bf16* __restrict__ dzt, float* __restrict__ part, int X,
               int Y, int Z, int64_t V, int vbpc)
{
    constexpr int Cout = RT * 16, CP = CT * 16;
    constexpr int SD = Cout + 2, SX = CP + 2; // LDS row strides (elements)
    constexpr int NCOL = 3 * CT, CPW = (NCOL + 3) / 4;
    __shared__ bf16 sdz[64 * SD];
    __shared__ bf16 sx[66 * SX];
    __shared__ int szc[64];
    __shared__ int sok[64];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int i = lane & 15, g = lane >> 4;
    const int tg = blockIdx.y;
    const int dx = tg / 3 - 1, dy = tg % 3 - 1;
    const int64_t off = ((int64_t)dx * Y + dy) * Z;
    f32x4 acc[RT][CPW];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int c = 0; c < CPW; ++c)
            acc[r][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int64_t nvb = (V + 63) / 64;
    const int64_t vb0 = (int64_t)blockIdx.x * vbpc;
    for (int64_t vb = vb0; vb < vb0 + vbpc && vb < nvb; ++vb)
    {
        const int64_t v0 = vb * 64;
        // output gradient rows v0 .. v0+63, two channels per step
        for (int idx = tid; idx < 64 * (Cout / 2); idx += kBlock)
        {
            const int row = idx / (Cout / 2), cp = idx % (Cout / 2);
            uint32_t w = 0;
            if (v0 + row < V)
                w = *reinterpret_cast<const uint32_t*>(dzt + (v0 + row) * Cout +
                                                       2 * cp);
            *reinterpret_cast<uint32_t*>(sdz + row * SD + 2 * cp) = w;
        }
        // input rows s = v0 - 1 + r + off, r in [0, 66)
        for (int idx = tid; idx < 66 * (CP / 2); idx += kBlock)
        {
            const int r = idx / (CP / 2), cp = idx % (CP / 2);
            const int64_t s = v0 - 1 + r + off;
            uint32_t w = 0;
            if (s >= 0 && s < V && 2 * cp < cinp)
                w = *reinterpret_cast<const uint32_t*>(x + s * cinp + 2 * cp);
            *reinterpret_cast<uint32_t*>(sx + r * SX + 2 * cp) = w;
        }
        if (tid < 64)
        {
            const int64_t v = v0 + tid;
            int ok = 0, zc = 0;
            if (v < V)
            {
                zc = (int)(v % Z);
                const int64_t t = v / Z;
                const int yc = (int)(t % Y), xc = (int)(t / Y);
                ok = (unsigned)(xc + dx) < (unsigned)X &&
                     (unsigned)(yc + dy) < (unsigned)Y;
            }
            szc[tid] = zc;
            sok[tid] = ok;
        }
        __syncthreads();
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
        {
            const int k0 = kb * 32 + 8 * g;
            bf16x8 A[RT];
#pragma unroll
            for (int rt = 0; rt < RT; ++rt)
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    A[rt][e] = sdz[(k0 + e) * SD + rt * 16 + i];
            // validity of (voxel k0 + e, dz) as three 8-bit masks
            uint32_t okm[3] = {0u, 0u, 0u};
#pragma unroll
            for (int e = 0; e < 8; ++e)
            {
                const int zc = szc[k0 + e];
                const bool ok = sok[k0 + e] != 0;
#pragma unroll
                for (int d = 0; d < 3; ++d)
                    okm[d] |= (uint32_t)(ok && (unsigned)(zc + d - 1) <
                                                   (unsigned)Z)
                              << e;
            }
#pragma unroll
            for (int cw = 0; cw < CPW; ++cw)
            {
                const int col = wv + 4 * cw;
                if (col < NCOL)
                {
                    const int d = col / CT, ct = col % CT;
                    bf16x8 B;
#pragma unroll
                    for (int e = 0; e < 8; ++e)
                    {
                        const bf16 val =
                            sx[(k0 + e + d) * SX + ct * 16 + i];
                        B[e] = ((okm[d] >> e) & 1u) ? val : (bf16)0.0f;
                    }
#pragma unroll
                    for (int rt = 0; rt < RT; ++rt)
                        acc[rt][cw] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            A[rt], B, acc[rt][cw], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // lane holds D[4g + r][i]: output channel rt*16 + 4g + r, input channel
    // ct*16 + i
#pragma unroll
    for (int cw = 0; cw < CPW; ++cw)
    {
        const int col = wv + 4 * cw;
        if (col >= NCOL)
            continue;
        const int d = col / CT, ct = col % CT;
        const int tap = tg * 3 + d;
        const int ci = ct * 16 + i;
        if (ci >= cinp)
            continue;
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int r = 0; r < 4; ++r)
            {
                const int co = rt * 16 + 4 * g + r;
                part[(((int64_t)blockIdx.x * 27 + tap) * cinp + ci) * Cout + co] =
                    acc[rt][cw][r];
            }
    }
}

// grad[(tap * cin + ci) * cout + co] += sum over chunks of the partials
__global__ void __launch_bounds__(kBlock)
    wgrad_reduce(const float* __restrict__ part, int nchunks, int cin,
                 int cinp, int cout, float* __restrict__ grad)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = 27LL * cin * cout;
    if (t >= n)
        return;
    const int co = (int)(t % cout);
    const int64_t r = t / cout;
    const int ci = (int)(r % cin), tap = (int)(r / cin);
    float s = 0.0f;
    for (int c = 0; c < nchunks; ++c)
        s += part[(((int64_t)c * 27 + tap) * cinp + ci) * cout + co];
    grad[t] += s;
}

// per-block column sums of a [V][C] BF16 tensor: part[block][c]
__global__ void __launch_bounds__(kBlock)
    colsum_part(const bf16* __restrict__ a, int C, int64_t V, int64_t vpb,
                float* __restrict__ part)
{
    __shared__ float red[kBlock];
    const int c = threadIdx.x % C;
    const int lanes = kBlock / C; // threads per channel
    const int j = threadIdx.x / C;
    float s = 0.0f;
    if (j < lanes)
    {
        const int64_t v0 = (int64_t)blockIdx.x * vpb;
        for (int64_t v = v0 + j; v < v0 + vpb && v < V; v += lanes)
            s += (float)a[v * C + c];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < C)
    {
        float t = 0.0f;
        for (int q = 0; q < lanes; ++q)
            t += red[q * C + threadIdx.x];
        part[(int64_t)blockIdx.x * C + threadIdx.x] = t;
    }
}

// dst[c] += sum_b part[b][c]
__global__ void colsum_reduce(const float* __restrict__ part, int nb, int C,
                              float* __restrict__ dst)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C)
        return;
    float s = 0.0f;
    for (int b = 0; b < nb; ++b)
        s += part[(int64_t)b * C + c];
    dst[c] += s;
}

// ------------------------------------------------------------ elementwise
// The work-items of one dispatch along x are counted in 32 bits, so a launch
// of one thread per (voxel, channel) of a 512^3 activation (up to 48 x 2^27
// elements) cannot cover its range: these kernels loop over their range with
// the grid's stride and are launched with at most kMaxGrid workgroups.
#define VC_GRID_STRIDE(t, n)                                                  \
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (n); \
         t += (int64_t)gridDim.x * blockDim.x)
__global__ void __launch_bounds__(kBlock)
    relu_mask(bf16* __restrict__ g, const bf16* __restrict__ y, int64_t n)
{
    VC_GRID_STRIDE(t, n)
    if (!((float)y[t] > 0.0f))
        g[t] = (bf16)0.0f;
}

// child voxel (x, y, z) of level dims (X, Y, Z) -> parent at (X/2, Y/2, Z/2)
__device__ __forceinline__ int64_t parent_of(int64_t v, int X, int Y, int Z)
{
    const int z = (int)(v % Z);
    const int64_t r = v / Z;
    const int y = (int)(r % Y), x = (int)(r / Y);
    return (((int64_t)(x >> 1) * (Y >> 1)) + (y >> 1)) * (Z >> 1) + (z >> 1);
}

// p[parent][c] = mean of the 8 children (dims of the child level X, Y, Z)
__global__ void __launch_bounds__(kBlock)
    pool_fwd(const bf16* __restrict__ a, bf16* __restrict__ p, int C, int X,
             int Y, int Z)
{
    // one thread per 8 channels of a pooled voxel (C is a multiple of 8):
    // 16-byte loads and stores; each channel is summed in the cell order
    // (x, y, z lexicographic) from 0, then scaled by 1/8
    const int Xp = X >> 1, Yp = Y >> 1, Zp = Z >> 1, C8 = C >> 3;
    const int64_t n = (int64_t)Xp * Yp * Zp * C8;
    VC_GRID_STRIDE(t, n)
    {
        const int c8 = (int)(t % C8);
        const int64_t v = t / C8;
        const int z = (int)(v % Zp);
        const int64_t r = v / Zp;
        const int y = (int)(r % Yp), x = (int)(r / Yp);
        float s[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            s[k] = 0.0f;
#pragma unroll
        for (int q = 0; q < 8; ++q)
        {
            const int64_t cv =
                (((int64_t)(2 * x + (q >> 2)) * Y) + (2 * y + ((q >> 1) & 1))) * Z +
                (2 * z + (q & 1));
            const bf16x8 w = *reinterpret_cast<const bf16x8*>(a + cv * C + 8 * c8);
#pragma unroll
            for (int k = 0; k < 8; ++k)
                s[k] += (float)w[k];
        }
        bf16x8 o;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            o[k] = (bf16)(s[k] * 0.125f);
        *reinterpret_cast<bf16x8*>(p + v * C + 8 * c8) = o;
    }
}

// g_child += g_parent / 8
__global__ void __launch_bounds__(kBlock)
    pool_bwd_acc(bf16* __restrict__ ga, const bf16* __restrict__ gp, int C,
                 int X, int Y, int Z)
{
    const int64_t n = (int64_t)X * Y * Z * C;
    VC_GRID_STRIDE(t, n)
    {
        const int c = (int)(t % C);
        const int64_t v = t / C;
        const int64_t pv = parent_of(v, X, Y, Z);
        ga[t] = (bf16)((float)ga[t] + 0.125f * (float)gp[pv * C + c]);
    }
}

// cat[v] = [up[parent(v)] (C1 channels), skip[v] (C2)] at child dims
__global__ void __launch_bounds__(kBlock)
    concat_fwd(const bf16* __restrict__ up, int C1, const bf16* __restrict__ sk,
               int C2, bf16* __restrict__ cat, int X, int Y, int Z)
{
    const int C = C1 + C2;
    const int64_t n = (int64_t)X * Y * Z * C;
    VC_GRID_STRIDE(t, n)
    {
        const int c = (int)(t % C);
        const int64_t v = t / C;
        cat[t] = c < C1 ? up[parent_of(v, X, Y, Z) * C1 + c]
                        : sk[v * C2 + (c - C1)];
    }
}

// g_up[parent][c] = sum over the 8 children of g_cat[child][c] (c < C1)
__global__ void __launch_bounds__(kBlock)
    split_up(const bf16* __restrict__ gcat, int C, int C1,
             bf16* __restrict__ gup, int X, int Y, int Z)
{
    const int Xp = X >> 1, Yp = Y >> 1, Zp = Z >> 1;
    const int64_t n = (int64_t)Xp * Yp * Zp * C1;
    VC_GRID_STRIDE(t, n)
    {
        const int c = (int)(t % C1);
        const int64_t v = t / C1;
        const int z = (int)(v % Zp);
        const int64_t r = v / Zp;
        const int y = (int)(r % Yp), x = (int)(r / Yp);
        float s = 0.0f;
#pragma unroll
        for (int q = 0; q < 8; ++q)
        {
            const int64_t cv =
                (((int64_t)(2 * x + (q >> 2)) * Y) + (2 * y + ((q >> 1) & 1))) * Z +
                (2 * z + (q & 1));
            s += (float)gcat[cv * C + c];
        }
        gup[t] = (bf16)s;
    }
}

// g_skip[v][c] = g_cat[v][C1 + c]
__global__ void __launch_bounds__(kBlock)
    split_skip(const bf16* __restrict__ gcat, int C, int C1,
               bf16* __restrict__ gsk, int64_t V)
{
    const int C2 = C - C1;
    VC_GRID_STRIDE(t, V * C2)
    {
        const int c = (int)(t % C2);
        const int64_t v = t / C2;
        gsk[t] = gcat[v * C + C1 + c];
    }
}

// network input over a region of dims (X, Y, Z) whose voxel (0,0,0) is
// volume voxel (ox, oy, oz): x0 = [occ, sqrt(max(edt, 0)) / 32] (the first
// layer's input, 4 bytes per voxel) and, for training, x8 = the same
// padded to 8 channels (the layout of layer 0's weight gradient)
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__global__ void __launch_bounds__(kBlock)
    prep_input(const uint8_t* __restrict__ occ, const int32_t* __restrict__ edt,
               int nx, int ny, int nz, int ox, int oy, int oz,
               bf16* __restrict__ x0, bf16* __restrict__ x8, int X, int Y, int Z)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t V = (int64_t)X * Y * Z;
    if (v >= V)
        return;
    const int z = (int)(v % Z);
    const int64_t r = v / Z;
    const int y = (int)(r % Y), x = (int)(r / Y);
    const int gx = ox + x, gy = oy + y, gz = oz + z;
    float o = 0.0f, d = 0.0f;
    if ((unsigned)gx < (unsigned)nx && (unsigned)gy < (unsigned)ny &&
        (unsigned)gz < (unsigned)nz)
    {
        const int64_t i = ((int64_t)gx * ny + gy) * nz + gz;
        o = occ[i] ? 1.0f : 0.0f;
        const int32_t e = edt[i];
        d = e > 0 ? sqrtf((float)e) / kScale : 0.0f;
    }
    const bf16x2 w2 = {(bf16)o, (bf16)d};
    reinterpret_cast<bf16x2*>(x0)[v] = w2;
    if (x8)
    {
        bf16x8 w = {};
        w[0] = w2[0];
        w[1] = w2[1];
        reinterpret_cast<bf16x8*>(x8)[v] = w;
    }
}

// head: out[v][o] = sum_c d0[v][c] Wh[c][o] + bh[o] (FP32)
__global__ void __launch_bounds__(kBlock)
    head_fwd(const bf16* __restrict__ d0, const float* __restrict__ Wh,
             float* __restrict__ out, int64_t V)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V)
        return;
    float o0 = Wh[kHeadIn * kHeadOut], o1 = Wh[kHeadIn * kHeadOut + 1];
#pragma unroll
    for (int c = 0; c < kHeadIn; ++c)
    {
        const float h = (float)d0[v * kHeadIn + c];
        o0 += h * Wh[c * 2];
        o1 += h * Wh[c * 2 + 1];
    }
    out[2 * v] = o0;
    out[2 * v + 1] = o1;
}

// loss terms and dL/d(out) on inside voxels of the region (volume voxel =
// region voxel for training: no offset); part[block] = block loss sum
__global__ void __launch_bounds__(kBlock)
    loss_bwd(const float* __restrict__ out, const uint8_t* __restrict__ occ,
             const uint32_t* __restrict__ lam2, int nx, int ny, int nz,
             int X, int Y, int Z, const unsigned long long* __restrict__ cnt,
             float* __restrict__ dout, float* __restrict__ part)
{
    __shared__ float red[kBlock];
    const unsigned long long nin = *cnt;
    const float inv_n = 1.0f / (float)(nin ? nin : 1ull);
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t V = (int64_t)X * Y * Z;
    float l = 0.0f;
    if (v < V)
    {
        const int z = (int)(v % Z);
        const int64_t r = v / Z;
        const int y = (int)(r % Y), x = (int)(r / Y);
        float g0 = 0.0f, g1 = 0.0f;
        if (x < nx && y < ny && z < nz)
        {
            const int64_t i = ((int64_t)x * ny + y) * nz + z;
            if (occ[i])
            {
                const float t = sqrtf((float)lam2[i]) / kScale;
                const float den = t + kEpsT;
                const float rr = (out[2 * v] - t) / den;
                const float e = logf(fabsf(rr) + kErrFloor);
                const float q = out[2 * v + 1] - e;
                l = rr * rr + kErrW * q * q;
                g0 = 2.0f * rr / den * inv_n;
                g1 = 2.0f * kErrW * q * inv_n;
            }
        }
        dout[2 * v] = g0;
        dout[2 * v + 1] = g1;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1)
    {
        if (threadIdx.x < s)
            red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        part[blockIdx.x] = red[0];
}

// head backward: gd0[v][c] = sum_o dout[v][o] Wh[c][o]; per-block partials
// of dWh[c][o] (32 values) and dbh[o] (2) in part[block][34]
__global__ void __launch_bounds__(kBlock)
    head_bwd(const bf16* __restrict__ d0, const float* __restrict__ dout,
             const float* __restrict__ Wh, bf16* __restrict__ gd0, int64_t V,
             float* __restrict__ part)
{
    constexpr int NP = kHeadIn * kHeadOut + kHeadOut;
    __shared__ float red[NP][kBlock / 64];
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float g0 = 0.0f, g1 = 0.0f;
    float h[kHeadIn];
    if (v < V)
    {
        g0 = dout[2 * v];
        g1 = dout[2 * v + 1];
    }
#pragma unroll
    for (int c = 0; c < kHeadIn; ++c)
    {
        h[c] = v < V ? (float)d0[v * kHeadIn + c] : 0.0f;
        if (v < V)
            gd0[v * kHeadIn + c] = (bf16)(g0 * Wh[c * 2] + g1 * Wh[c * 2 + 1]);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int p = 0; p < NP; ++p)
    {
        float val = p < kHeadIn * kHeadOut
                        ? h[p >> 1] * ((p & 1) ? g1 : g0)
                        : ((p - kHeadIn * kHeadOut) ? g1 : g0);
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1)
            val += __shfl_xor(val, d);
        if (lane == 0)
            red[p][wv] = val;
    }
    __syncthreads();
    if (threadIdx.x < NP)
    {
        float s = 0.0f;
        for (int w = 0; w < kBlock / 64; ++w)
            s += red[threadIdx.x][w];
        part[(int64_t)blockIdx.x * NP + threadIdx.x] = s;
    }
}

__global__ void __launch_bounds__(kBlock)
    adam_step(float* __restrict__ w, const float* __restrict__ g,
              float* __restrict__ m, float* __restrict__ s, int64_t n,
              float lr, float gscale, float bc1, float bc2)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    const float gg = g[t] * gscale;
    const float mm = 0.9f * m[t] + 0.1f * gg;
    const float ss = 0.999f * s[t] + 0.001f * gg * gg;
    m[t] = mm;
    s[t] = ss;
    w[t] -= lr * (mm / bc1) / (sqrtf(ss / bc2) + 1e-8f);
}

// The first layer (2 input channels, 16 outputs): a K pair is 4 taps x 2
// channels (taps 4p .. 4p + 3) instead of one tap's 8-channel chunk with 6
// zero channels, so a 16-voxel tile takes 2 MFMAs (7 pairs) instead of 7
// (27 pairs), and the brick's input is staged as its 2 channels, 4 bytes per
// voxel.  Weights: W4[pair][cout][8], pair p slot e = tap 4p + e/2, channel
// e % 2 (zero past tap 26).  Bias + ReLU + zeroing outside the padded
// volume as conv_fwd_lds.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
template <int BX, int BY, int NZT>
__global__ void __launch_bounds__(kBlock)
    conv_first_lds(const bf16* __restrict__ in, bf16* __restrict__ out,
                   const bf16x8* __restrict__ W4, const float* __restrict__ bias, int X,
                   int Y, int Z, Box box, int off32)
{
    constexpr int MT = BX * BY * NZT / 4;
    static_assert(MT * 4 == BX * BY * NZT, "tiles per wavefront");
    constexpr int BZ = 16 * NZT, RY = BY + 2, RZ = BZ + 2;
    constexpr int NR = (BX + 2) * RY * RZ;
    __shared__ uint32_t reg2[NR]; // channels 0 and 1 of each region voxel
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = lane & 15, g = lane >> 4;
    const int nbz = (Z + BZ - 1) / BZ, nby = (Y + BY - 1) / BY;
    const int nbx = (X + BX - 1) / BX;
    int blk = conv_brick(nbz * nby * nbx);
    if (blk < 0)
        return;
    const int bz = blk % nbz;
    blk /= nbz;
    // x groups of VC_CONV_GX bricks, y within a group, then the groups
    const int q = blk / (nby * VC_CONV_GX), rem = blk % (nby * VC_CONV_GX);
    const int gsz = min(VC_CONV_GX, nbx - q * VC_CONV_GX);
    const int by = rem / gsz, bx = q * VC_CONV_GX + rem % gsz;
    const int x0 = BX * bx, y0 = BY * by, z0 = bz * BZ;
    int tb[MT];
    f32x4 acc[MT][1];
#pragma unroll
    for (int m = 0; m < MT; ++m)
    {
        const int t = wv * MT + m, row = t / NZT, seg = t % NZT;
        tb[m] = ((row / BY + 1) * RY + (row % BY + 1)) * RZ + 1 + 16 * seg;
        acc[m][0] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    // the region (a brick inside the volume: no bounds checks, 32-bit offsets
    // from its first voxel, every load issued before the LDS stores)
    constexpr int NIT = (NR + kBlock - 1) / kBlock;
    if (off32 && x0 >= 1 && y0 >= 1 && z0 >= 1 && x0 + BX < X && y0 + BY < Y &&
        z0 + BZ < Z)
    {
        const uint32_t* base = reinterpret_cast<const uint32_t*>(in) +
                               (((int64_t)(x0 - 1)) * Y + (y0 - 1)) * Z + (z0 - 1);
        uint32_t v[NIT];
#pragma unroll
        for (int k = 0; k < NIT; ++k)
        {
            const int e = min((int)threadIdx.x + k * kBlock, NR - 1);
            const int rz = e % RZ, rxy = e / RZ;
            v[k] = base[((rxy / RY) * Y + rxy % RY) * Z + rz];
        }
#pragma unroll
        for (int k = 0; k < NIT; ++k)
        {
            const int e = threadIdx.x + k * kBlock;
            if (k < NIT - 1 || e < NR)
                reg2[e] = v[k];
        }
    }
    else
        for (int e = threadIdx.x; e < NR; e += kBlock)
        {
            const int rz = e % RZ, rxy = e / RZ;
            const int xx = x0 - 1 + rxy / RY, yy = y0 - 1 + rxy % RY, zz = z0 - 1 + rz;
            uint32_t v = 0u;
            if ((unsigned)xx < (unsigned)X && (unsigned)yy < (unsigned)Y &&
                (unsigned)zz < (unsigned)Z)
                v = reinterpret_cast<const uint32_t*>(in)[((int64_t)xx * Y + yy) * Z + zz];
            reg2[e] = v;
        }
    __syncthreads();
#pragma unroll
    for (int p0 = 0; p0 < 8; p0 += 4)
    {
        const int p = p0 + g; // taps 4p .. 4p + 3 (pair 7 is all zero)
        int ro[4];
        bool tv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            const int tap = 4 * p + j;
            tv[j] = tap < 27;
            const int tt = tv[j] ? tap : 13;
            ro[j] = ((tt / 9 - 1) * RY + (tt / 3) % 3 - 1) * RZ + tt % 3 - 1 + i;
        }
        const bf16x8 b = W4[p * 16 + i];
#pragma unroll
        for (int m = 0; m < MT; ++m)
        {
            const u32x4 w = {tv[0] ? reg2[ro[0] + tb[m]] : 0u, tv[1] ? reg2[ro[1] + tb[m]] : 0u,
                             tv[2] ? reg2[ro[2] + tb[m]] : 0u, tv[3] ? reg2[ro[3] + tb[m]] : 0u};
            acc[m][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                b, __builtin_bit_cast(bf16x8, w), acc[m][0], 0, 0, 0);
        }
    }
    conv_store<1, MT, BY, NZT, true>(acc, out, bias, x0, y0, z0, X, Y, Z, box, off32);
}

// Inference: the first two layers in one kernel.  Layer 1's input region
// (6 x 6 x 18 voxels around a 4 x 4 x 16 brick) is layer 0's output there,
// computed into LDS from the 2-channel input over an 8 x 8 x 20 region instead
// of being stored by one kernel and staged back by the next.  Layer 0's values
// are those of conv_first_lds bit for bit (the same K pairs, MFMA order, bias,
// ReLU, BF16 rounding and zeroing outside the padded box; region voxels
// outside the volume are 0 as layer 1's staging reads them), so layer 1's
// output equals the unfused pair's.  Layer 0's 648 region voxels form 41
// tiles of 16 (the last one half used), spread over the 4 wavefronts.
// The network input (prep_input's two channels) is computed from the volume's
// occupancy and EDT as the region is staged, so it is never stored either.
template <int BX, int BY>
__global__ void __launch_bounds__(kBlock)
    conv01_lds(const uint8_t* __restrict__ occ, const int32_t* __restrict__ edt, int nx,
               int ny, int nz, bf16* __restrict__ out, const bf16x8* __restrict__ W4,
               const float* __restrict__ bias0, const bf16x8* __restrict__ W1,
               const float* __restrict__ bias1, int X, int Y, int Z, Box box, int off32)
{
    constexpr int NZT = 1, MT = BX * BY / 4, SC = 2;
    static_assert(MT == BY, "a wavefront holds one brick x row");
    constexpr int BZ = 16, RY = BY + 2, RZ = BZ + 2, NR = (BX + 2) * RY * RZ;
    constexpr int IY = BY + 4, IZ = BZ + 4, NI = (BX + 4) * IY * IZ; // layer 0's input
    constexpr int NT0 = (NR + 15) / 16;                               // layer 0 tiles
    __shared__ uint32_t in2[NI];
    __shared__ bf16x8 reg[SC * NR];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = lane & 15, g = lane >> 4;
    const int nbz = (Z + BZ - 1) / BZ, nby = (Y + BY - 1) / BY;
    const int nbx = (X + BX - 1) / BX;
    int blk = conv_brick(nbz * nby * nbx);
    if (blk < 0)
        return;
    const int bz = blk % nbz;
    blk /= nbz;
    const int q = blk / (nby * VC_CONV_GX), rem = blk % (nby * VC_CONV_GX);
    const int gsz = min(VC_CONV_GX, nbx - q * VC_CONV_GX);
    const int by = rem / gsz, bx = q * VC_CONV_GX + rem % gsz;
    const int x0 = BX * bx, y0 = BY * by, z0 = bz * BZ;
    // layer 0's input region: voxels x0 - 2 .. x0 + BX + 1, ... of the padded
    // region, whose voxel (0, 0, 0) is volume voxel (box.ox, box.oy, box.oz):
    // [occ, sqrt(max(edt, 0)) / 32] as prep_input computes it, 0 outside
    for (int e = threadIdx.x; e < NI; e += kBlock)
    {
        const int iz = e % IZ, ixy = e / IZ;
        const int xx = x0 - 2 + ixy / IY, yy = y0 - 2 + ixy % IY, zz = z0 - 2 + iz;
        const int gx = box.ox + xx, gy = box.oy + yy, gz = box.oz + zz;
        uint32_t v = 0u;
        if ((unsigned)xx < (unsigned)X && (unsigned)yy < (unsigned)Y &&
            (unsigned)zz < (unsigned)Z && (unsigned)gx < (unsigned)nx &&
            (unsigned)gy < (unsigned)ny && (unsigned)gz < (unsigned)nz)
        {
            const int64_t gi = ((int64_t)gx * ny + gy) * nz + gz;
            const float o = occ[gi] ? 1.0f : 0.0f;
            const int32_t ed = edt[gi];
            const float dd = ed > 0 ? sqrtf((float)ed) / kScale : 0.0f;
            const bf16x2 w2 = {(bf16)o, (bf16)dd};
            v = __builtin_bit_cast(uint32_t, w2);
        }
        in2[e] = v;
    }
    __syncthreads();
    // layer 0 over layer 1's region (voxel r = (rx RY + ry) RZ + rz)
    {
        float bv[4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
            bv[r] = bias0[4 * g + r];
        bf16x8 wb[2];
        int to[2][4];
        bool tv[2][4];
#pragma unroll
        for (int h = 0; h < 2; ++h)
        {
            const int p = 4 * h + g; // taps 4p .. 4p + 3 (pair 7 is all zero)
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                const int tap = 4 * p + j;
                tv[h][j] = tap < 27;
                const int tt = tv[h][j] ? tap : 13;
                to[h][j] = ((tt / 9 - 1) * IY + (tt / 3) % 3 - 1) * IZ + tt % 3 - 1;
            }
            wb[h] = W4[p * 16 + i];
        }
        for (int t = wv; t < NT0; t += kBlock / 64)
        {
            const int r = min(16 * t + i, NR - 1);
            const int rz = r % RZ, rxy = r / RZ, rx = rxy / RY, ry = rxy % RY;
            // the voxel's centre in layer 0's input region
            const int ic = ((rx + 1) * IY + (ry + 1)) * IZ + rz + 1;
            f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int h = 0; h < 2; ++h)
            {
                const u32x4 w = {tv[h][0] ? in2[ic + to[h][0]] : 0u,
                                 tv[h][1] ? in2[ic + to[h][1]] : 0u,
                                 tv[h][2] ? in2[ic + to[h][2]] : 0u,
                                 tv[h][3] ? in2[ic + to[h][3]] : 0u};
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wb[h], __builtin_bit_cast(bf16x8, w),
                                                              acc, 0, 0, 0);
            }
            // lane (i, g): channels 4 g .. 4 g + 3 of voxel 16 t + i
            const int x = x0 - 1 + rx, y = y0 - 1 + ry, z = z0 - 1 + rz;
            const bool invol = (unsigned)x < (unsigned)X && (unsigned)y < (unsigned)Y &&
                               (unsigned)z < (unsigned)Z;
            const bool inbox = (unsigned)(box.ox + x) < (unsigned)box.px &&
                               (unsigned)(box.oy + y) < (unsigned)box.py &&
                               (unsigned)(box.oz + z) < (unsigned)box.pz;
            f32x4 val = acc + f32x4{bv[0], bv[1], bv[2], bv[3]};
#pragma unroll
            for (int k = 0; k < 4; ++k)
                val[k] = fmaxf(val[k], 0.0f);
            u32x2 qv = __builtin_bit_cast(u32x2, __builtin_convertvector(val, bf16x4));
            qv = invol && inbox ? qv : u32x2{0u, 0u};
            if (16 * t + i < NR)
                reinterpret_cast<u32x2*>(reg + (g >> 1) * NR + r)[g & 1] = qv;
        }
    }
    __syncthreads();
    // layer 1 (16 -> 16) over the staged region
    int tb[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
        tb[m] = ((wv + 1) * RY + (m + 1)) * RZ + 1;
    f32x4 acc[MT][1];
#pragma unroll
    for (int m = 0; m < MT; ++m)
        acc[m][0] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    conv_slice<1, MT, NR, RY, RZ, SC>(reg, tb, W1, 0, 2, acc);
    conv_store<1, MT, BY, NZT, true>(acc, out, bias1, x0, y0, z0, X, Y, Z, box, off32);
}

// first layer's W[27][2][16] -> W4[8][16][8] (pair p slot e: tap 4p + e/2,
// channel e % 2)
__global__ void __launch_bounds__(kBlock)
    pack_first(const float* __restrict__ W, bf16* __restrict__ W4)
{
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= 8 * 16 * 8)
        return;
    const int e = t & 7, co = (t >> 3) & 15, p = t >> 7;
    const int tap = 4 * p + (e >> 1), ch = e & 1;
    W4[t] = (bf16)(tap < 27 ? W[(tap * 2 + ch) * 16 + co] : 0.0f);
}

// FP32 master W[27][cin][cout] -> forward chunks Wf[kc][cout][8] and the
// flipped, transposed input-gradient chunks Wd[kcd][cinp][8]
__global__ void __launch_bounds__(kBlock)
    pack_fwd(const float* __restrict__ W, int cin, int cinp, int cout, int kc,
             bf16* __restrict__ Wf)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)kc * cout * 8)
        return;
    const int e = (int)(t & 7);
    const int co = (int)((t >> 3) % cout);
    const int k = (int)((t >> 3) / cout);
    const int cin8 = cinp >> 3;
    float val = 0.0f;
    if (k < 27 * cin8)
    {
        const int tap = k / cin8, ci = (k % cin8) * 8 + e;
        if (ci < cin)
            val = W[((int64_t)tap * cin + ci) * cout + co];
    }
    Wf[t] = (bf16)val;
}

__global__ void __launch_bounds__(kBlock)
    pack_dgrad(const float* __restrict__ W, int cin, int cinp, int cout,
               int kcd, bf16* __restrict__ Wd)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)kcd * cinp * 8)
        return;
    const int e = (int)(t & 7);
    const int ci = (int)((t >> 3) % cinp);
    const int k = (int)((t >> 3) / cinp);
    const int cout8 = cout >> 3;
    float val = 0.0f;
    if (k < 27 * cout8 && ci < cin)
    {
        const int tap = 26 - k / cout8, co = (k % cout8) * 8 + e;
        val = W[((int64_t)tap * cin + ci) * cout + co];
    }
    Wd[t] = (bf16)val;
}

__global__ void __launch_bounds__(kBlock)
    fill_f32(float* __restrict__ p, int64_t n, float v)
{
    VC_GRID_STRIDE(t, n)
    p[t] = v;
}

__global__ void __launch_bounds__(kBlock)
    bf16_to_f32(const bf16* __restrict__ a, float* __restrict__ b, int64_t n)
{
    VC_GRID_STRIDE(t, n)
    b[t] = (float)a[t];
}

// core of a region's head output -> volume outputs (lambda, relative error)
__global__ void __launch_bounds__(kBlock)
    copy_core(const float* __restrict__ out, int X, int Y, int Z, int ox,
              int oy, int oz, int cx0, int cy0, int cz0, int cx1, int cy1,
              int cz1, float* __restrict__ lam, float* __restrict__ rel,
              int nx, int ny, int nz)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t V = (int64_t)X * Y * Z;
    if (v >= V)
        return;
    const int z = (int)(v % Z);
    const int64_t r = v / Z;
    const int y = (int)(r % Y), x = (int)(r / Y);
    const int gx = ox + x, gy = oy + y, gz = oz + z;
    if (gx < cx0 || gx >= cx1 || gy < cy0 || gy >= cy1 || gz < cz0 ||
        gz >= cz1)
        return;
    if (gx >= nx || gy >= ny || gz >= nz)
        return;
    const int64_t i = ((int64_t)gx * ny + gy) * nz + gz;
    lam[i] = out[2 * v] * kScale;
    rel[i] = expf(out[2 * v + 1]);
}

// gated measure: inside voxels keep the surrogate lambda when the predicted
// relative error is <= tau, else the exact sqrt(lam2); ctr[0] accepted,
// ctr[1] fallback, ctr[2] inside
__global__ void __launch_bounds__(kBlock)
    gate_count(const uint8_t* __restrict__ occ, const float* __restrict__ rel,
               float tau, int64_t N, unsigned long long* __restrict__ ctr)
{
    // a capped grid strides over the volume; counts are reduced per
    // workgroup, so each workgroup adds once per counter (one add per
    // wavefront and counter serialised millions of atomics on 3 addresses)
    __shared__ unsigned long long sh[2][kBlock / 64];
    unsigned long long acc = 0, fb = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * blockDim.x)
        if (occ[i])
        {
            if (rel[i] <= tau)
                ++acc;
            else
                ++fb;
        }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
    {
        acc += __shfl_xor(acc, d);
        fb += __shfl_xor(fb, d);
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
    {
        sh[0][wv] = acc;
        sh[1][wv] = fb;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned long long a = 0, f = 0;
        for (int w = 0; w < kBlock / 64; ++w)
        {
            a += sh[0][w];
            f += sh[1][w];
        }
        if (a | f)
        {
            atomicAdd(&ctr[0], a);
            atomicAdd(&ctr[1], f);
            atomicAdd(&ctr[2], a + f);
        }
    }
}

__global__ void __launch_bounds__(kBlock)
    gate_mix(const uint8_t* __restrict__ occ, const float* __restrict__ lam,
             const float* __restrict__ rel, const uint32_t* __restrict__ lam2,
             float tau, int64_t N, float* __restrict__ meas)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    float m = 0.0f;
    if (occ[i])
        m = rel[i] <= tau ? fmaxf(lam[i], 0.0f) : sqrtf((float)lam2[i]);
    meas[i] = m;
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }
// grid of a VC_GRID_STRIDE kernel: one thread per element up to kMaxGrid
// workgroups (2^28 threads), fewer trips of the loop otherwise
constexpr int64_t kMaxGrid = int64_t(1) << 20;
unsigned grid_for(int64_t n)
{
    const int64_t b = (n + kBlock - 1) / kBlock;
    return (unsigned)(b < kMaxGrid ? (b > 0 ? b : 1) : kMaxGrid);
}

} // namespace

// ------------------------------------------------------------ host side
struct vc_surrogate
{
    vc_ctx* ctx = nullptr;
    int64_t np = 0;
    float* theta = nullptr; // FP32 master
    float* grad = nullptr;
    float* m1 = nullptr;
    float* m2 = nullptr;
    int64_t step = 0;
    bf16* Wf[14] = {};
    bf16* Wd[14] = {};
    bf16* W4 = nullptr; // the first layer's tap-packed weights (conv_first_lds)
    int kc[14] = {}, kcd[14] = {};
    bool packed = false;
    // workspace for one region
    int X = 0, Y = 0, Z = 0;
    std::vector<void*> ws;
    bf16* act[14] = {}; // ReLU outputs of L0a..L6b
    bf16 *x0 = nullptr, *p1 = nullptr, *p2 = nullptr, *p3 = nullptr;
    bf16* x8 = nullptr; // the input padded to 8 channels (training)
    // whether act[0] and act[13] hold the last forward pass's values (an
    // inference computes layers 0 and 1 in one kernel and evaluates the head
    // inside layer 13's convolution, storing neither)
    bool full_acts = false;
    // the volume an inference reads (layers 0 and 1 in one kernel read its
    // occupancy and EDT directly)
    const uint8_t* inf_occ = nullptr;
    const int32_t* inf_edt = nullptr;
    int inf_n[3] = {0, 0, 0};
    // and where its head writes the core voxels' lam / rel (no copy_core)
    CoreOut inf_core{};
    bf16 *cat2 = nullptr, *cat1 = nullptr, *cat0 = nullptr;
    float* out = nullptr;
    // backward workspace (allocated on first use)
    bool bw = false;
    bf16* gact[14] = {};
    bf16 *gcat2 = nullptr, *gcat1 = nullptr, *gcat0 = nullptr;
    bf16 *gp1 = nullptr, *gp2 = nullptr, *gp3 = nullptr;
    float* dout = nullptr;
    float* part = nullptr;
    size_t part_floats = 0;
    // volume-sized outputs of inference
    float* vlam = nullptr;
    float* vrel = nullptr;
    float* vmeas = nullptr;
    int64_t vN = 0;
    int org[3] = {0, 0, 0}; // region origin in the padded volume
    int pad[3] = {0, 0, 0}; // padded volume dims (multiples of 8)
    unsigned long long* ctr = nullptr; // 64 counters
    float* lossbuf = nullptr;          // 64 per-volume losses
};

namespace
{

void free_ws(vc_surrogate* s)
{
    for (void* p : s->ws)
        vc_free(p);
    s->ws.clear();
    for (int l = 0; l < 14; ++l)
    {
        s->act[l] = nullptr;
        s->gact[l] = nullptr;
    }
    s->x0 = s->x8 = s->p1 = s->p2 = s->p3 = s->cat2 = s->cat1 = s->cat0 = nullptr;
    s->gcat2 = s->gcat1 = s->gcat0 = s->gp1 = s->gp2 = s->gp3 = nullptr;
    s->out = s->dout = s->part = nullptr;
    s->part_floats = 0;
    s->bw = false;
    s->X = s->Y = s->Z = 0;
}

int ws_alloc(vc_ctx* ctx, vc_surrogate* s, void** p, size_t bytes)
{
    int rc = vc_alloc(ctx, p, bytes);
    if (rc == VC_OK)
        s->ws.push_back(*p);
    return rc;
}

int64_t level_voxels(const vc_surrogate* s, int level)
{
    return (int64_t)(s->X >> level) * (s->Y >> level) * (s->Z >> level);
}

// region origin (multiple of 8, may be negative) and padded-volume extent
// at level lv
Box level_box(const vc_surrogate* s, int lv)
{
    Box b;
    b.ox = s->org[0] / (1 << lv);
    b.oy = s->org[1] / (1 << lv);
    b.oz = s->org[2] / (1 << lv);
    b.px = s->pad[0] >> lv;
    b.py = s->pad[1] >> lv;
    b.pz = s->pad[2] >> lv;
    return b;
}

int ensure_ws(vc_ctx* ctx, vc_surrogate* s, int X, int Y, int Z, bool bw)
{
    if (s->X != X || s->Y != Y || s->Z != Z)
    {
        VC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        free_ws(s);
        s->X = X;
        s->Y = Y;
        s->Z = Z;
        int rc;
        auto A = [&](bf16** p, int level, int C) {
            return ws_alloc(ctx, s, (void**)p,
                            2 * (size_t)level_voxels(s, level) * C);
        };
        if ((rc = A(&s->x0, 0, 2)) != VC_OK)
            return rc;
        for (int l = 0; l < 14; ++l)
            if ((rc = A(&s->act[l], kLayers[l].level, kLayers[l].cout)) !=
                VC_OK)
                return rc;
        if ((rc = A(&s->p1, 1, 16)) != VC_OK || (rc = A(&s->p2, 2, 32)) != VC_OK ||
            (rc = A(&s->p3, 3, 48)) != VC_OK || (rc = A(&s->cat2, 2, 112)) != VC_OK ||
            (rc = A(&s->cat1, 1, 80)) != VC_OK || (rc = A(&s->cat0, 0, 48)) != VC_OK)
            return rc;
        if ((rc = ws_alloc(ctx, s, (void**)&s->out,
                           8 * (size_t)level_voxels(s, 0))) != VC_OK)
            return rc;
    }
    if (bw && !s->bw)
    {
        int rc;
        auto A = [&](bf16** p, int level, int C) {
            return ws_alloc(ctx, s, (void**)p,
                            2 * (size_t)level_voxels(s, level) * C);
        };
        for (int l = 1; l < 14; ++l)
            if ((rc = A(&s->gact[l], kLayers[l].level, kLayers[l].cout)) !=
                VC_OK)
                return rc;
        if ((rc = A(&s->gact[0], 0, 16)) != VC_OK || (rc = A(&s->x8, 0, 8)) != VC_OK ||
            (rc = A(&s->gp1, 1, 16)) != VC_OK || (rc = A(&s->gp2, 2, 32)) != VC_OK ||
            (rc = A(&s->gp3, 3, 48)) != VC_OK || (rc = A(&s->gcat2, 2, 112)) != VC_OK ||
            (rc = A(&s->gcat1, 1, 80)) != VC_OK || (rc = A(&s->gcat0, 0, 48)) != VC_OK)
            return rc;
        if ((rc = ws_alloc(ctx, s, (void**)&s->dout,
                           8 * (size_t)level_voxels(s, 0))) != VC_OK)
            return rc;
        // partials: weight gradients (chunks x 27 x cinp x cout, at most 256
        // chunks) and column / loss sums
        size_t pf = 0;
        for (int l = 0; l < 14; ++l)
        {
            const size_t f = 256ull * 27 * cinp_of(kLayers[l].cin) *
                             kLayers[l].cout;
            pf = f > pf ? f : pf;
        }
        const size_t cs = (size_t)(level_voxels(s, 0) / kBlock + 1) * 64;
        pf = cs > pf ? cs : pf;
        if ((rc = ws_alloc(ctx, s, (void**)&s->part, 4 * pf)) != VC_OK)
            return rc;
        s->part_floats = pf;
        s->bw = true;
    }
    return VC_OK;
}

int pack_weights(vc_ctx* ctx, vc_surrogate* s)
{
    hipStream_t S = ctx->stream;
    for (int l = 0; l < 14; ++l)
    {
        const LayerSpec& L = kLayers[l];
        const int cinp = cinp_of(L.cin);
        const float* W = s->theta + param_offset(l);
        const int64_t nf = (int64_t)s->kc[l] * L.cout * 8;
        pack_fwd<<<blocks_for(nf), kBlock, 0, S>>>(W, L.cin, cinp, L.cout,
                                                    s->kc[l], s->Wf[l]);
        VC_CHECK_LAUNCH(ctx);
        if (s->Wd[l])
        {
            const int64_t nd = (int64_t)s->kcd[l] * cinp * 8;
            pack_dgrad<<<blocks_for(nd), kBlock, 0, S>>>(
                W, L.cin, cinp, L.cout, s->kcd[l], s->Wd[l]);
            VC_CHECK_LAUNCH(ctx);
        }
    }
    pack_first<<<blocks_for(8 * 16 * 8), kBlock, 0, S>>>(s->theta + param_offset(0), s->W4);
    VC_CHECK_LAUNCH(ctx);
    s->packed = true;
    return VC_OK;
}

// brick of the LDS convolution: BX x BY rows, NZT tiles of 16 voxels in z,
// input staged SC chunks at a time
constexpr int kCBX = VC_CONV_BX, kCBY = VC_CONV_BY, kCNZT = VC_CONV_NZT, kCSC = VC_CONV_SC;

// whether a brick's input region (BX + 2 x planes) and output brick fit 31-bit
// element offsets from their first voxel
int conv_off32(int Y, int Z, int cmax)
{
    return (int64_t)(kCBX + 2) * Y * Z * cmax < (int64_t(1) << 31);
}

template <int NT, bool kCat, bool kRelu>
void launch_conv_t(hipStream_t S, const bf16* in, const bf16* in2, int c1, int cinp,
                   bf16* out, const bf16* W, const float* bias, int X, int Y, int Z,
                   const Box& box)
{
    const int64_t nb = (int64_t)((X + kCBX - 1) / kCBX) * ((Y + kCBY - 1) / kCBY) *
                       ((Z + 16 * kCNZT - 1) / (16 * kCNZT));
    conv_fwd_lds<NT, kCBX, kCBY, kCNZT, kCSC, kCat, kRelu><<<conv_grid(nb), kBlock, 0, S>>>(
        in, in2, c1, cinp, out, reinterpret_cast<const bf16x8*>(W), bias, X, Y, Z, box,
        conv_off32(Y, Z, std::max(cinp, NT * 16)));
}

// box: the region's origin and the padded-volume extent at this level.  With
// in2 the input is the concatenation [in (coarse level, c1 channels,
// upsampled) | in2 (cinp - c1 channels)].
int launch_conv(vc_ctx* ctx, const bf16* in, int cinp, bf16* out, int cout,
                const bf16* W, const float* bias, int X, int Y, int Z, int kc,
                int relu, const Box& b, const bf16* in2 = nullptr, int c1 = 0)
{
    hipStream_t S = ctx->stream;
    (void)kc;
    if (cinp % (8 * kCSC) != 0 || (in2 && c1 % 8 != 0))
        return vc_fail(ctx, VC_ERR_INVALID, "conv: %d input channels", cinp);
    // the decoder's concatenations are forward (ReLU) layers
    if (in2 && !relu)
        return vc_fail(ctx, VC_ERR_INVALID, "conv: concatenated input without ReLU");
#define VC_CONV_CASE(nt)                                                              \
    case nt:                                                                          \
        if (in2)                                                                      \
            launch_conv_t<nt, true, true>(S, in, in2, c1, cinp, out, W, bias, X, Y, Z, b); \
        else if (relu)                                                                \
            launch_conv_t<nt, false, true>(S, in, nullptr, 0, cinp, out, W, bias, X, Y, Z, \
                                           b);                                        \
        else                                                                          \
            launch_conv_t<nt, false, false>(S, in, nullptr, 0, cinp, out, W, bias, X, Y, \
                                            Z, b);                                    \
        break;
    switch (cout / 16)
    {
        VC_CONV_CASE(1)
        VC_CONV_CASE(2)
        VC_CONV_CASE(3)
        VC_CONV_CASE(4)
        VC_CONV_CASE(5)
        VC_CONV_CASE(7)
    default:
        return vc_fail(ctx, VC_ERR_INVALID, "conv: %d output channels", cout);
    }
#undef VC_CONV_CASE
    VC_CHECK_LAUNCH(ctx);
    return VC_OK;
}

template <int RT, int CT>
void launch_wgrad_t(hipStream_t S, const bf16* x, int cinp, const bf16* dz,
                    float* part, int X, int Y, int Z, int nchunks, int vbpc)
{
    const int64_t V = (int64_t)X * Y * Z;
    conv_wgrad<RT, CT><<<dim3((unsigned)nchunks, 9), kBlock, 0, S>>>(
        x, cinp, dz, part, X, Y, Z, V, vbpc);
}

// layer l: gradient of W (and b) from its input x and pre-activation
// gradient dz, accumulated into s->grad
int layer_wgrad(vc_ctx* ctx, vc_surrogate* s, int l, const bf16* x,
                const bf16* dz)
{
    const LayerSpec& L = kLayers[l];
    const int X = s->X >> L.level, Y = s->Y >> L.level, Z = s->Z >> L.level;
    const int64_t V = (int64_t)X * Y * Z;
    const int cinp = cinp_of(L.cin);
    const int64_t nvb = (V + 63) / 64;
    const int nchunks = (int)(nvb < 256 ? nvb : 256);
    const int vbpc = (int)((nvb + nchunks - 1) / nchunks);
    hipStream_t S = ctx->stream;
    const int RT = L.cout / 16, CT = (cinp + 15) / 16;
#define VC_WG(r, c)                                                            \
    if (RT == r && CT == c)                                                    \
        launch_wgrad_t<r, c>(S, x, cinp, dz, s->part, X, Y, Z, nchunks, vbpc); \
    else
    VC_WG(1, 1) VC_WG(2, 1) VC_WG(2, 2) VC_WG(3, 2) VC_WG(3, 3) VC_WG(4, 3)
    VC_WG(4, 4) VC_WG(3, 7) VC_WG(2, 5) VC_WG(1, 3)
    return vc_fail(ctx, VC_ERR_INVALID, "wgrad: no kernel for %d x %d", RT, CT);
#undef VC_WG
    VC_CHECK_LAUNCH(ctx);
    float* gW = s->grad + param_offset(l);
    const int64_t nw = 27LL * L.cin * L.cout;
    wgrad_reduce<<<blocks_for(nw), kBlock, 0, S>>>(s->part, nchunks, L.cin,
                                                    cinp, L.cout, gW);
    VC_CHECK_LAUNCH(ctx);
    // bias: column sums of dz
    const int64_t vpb = 4096;
    const int nb = (int)((V + vpb - 1) / vpb);
    colsum_part<<<nb, kBlock, 0, S>>>(dz, L.cout, V, vpb, s->part);
    VC_CHECK_LAUNCH(ctx);
    colsum_reduce<<<1, 64, 0, S>>>(s->part, nb, L.cout, gW + nw);
    VC_CHECK_LAUNCH(ctx);
    return VC_OK;
}

// forward over the current workspace region (input x0 filled); keep_cat
// materializes the decoder's concatenated inputs (the weight gradients read
// them), otherwise the first decoder convolution of a level reads the coarse
// level and the skip tensor directly
int forward(vc_ctx* ctx, vc_surrogate* s, bool keep_cat)
{
    hipStream_t S = ctx->stream;
    const bf16* in = s->x0;
    int rc;
    s->full_acts = keep_cat;
    auto conv = [&](int l, const bf16* x) {
        const LayerSpec& L = kLayers[l];
        if (l == 0)
        {
            const int64_t nb = (int64_t)((s->X + kCBX - 1) / kCBX) *
                               ((s->Y + kCBY - 1) / kCBY) *
                               ((s->Z + 16 * kCNZT - 1) / (16 * kCNZT));
            conv_first_lds<kCBX, kCBY, kCNZT><<<conv_grid(nb), kBlock, 0, S>>>(
                x, s->act[0], reinterpret_cast<const bf16x8*>(s->W4),
                s->theta + param_offset(0) + 27LL * L.cin * L.cout, s->X, s->Y, s->Z,
                level_box(s, 0), conv_off32(s->Y, s->Z, 16));
            VC_CHECK_LAUNCH(ctx);
            return VC_OK;
        }
        return launch_conv(ctx, x, cinp_of(L.cin), s->act[l], L.cout, s->Wf[l],
                           s->theta + param_offset(l) + 27LL * L.cin * L.cout,
                           s->X >> L.level, s->Y >> L.level, s->Z >> L.level,
                           s->kc[l], 1, level_box(s, L.level));
    };
    auto dims = [&](int lv, int& X, int& Y, int& Z) {
        X = s->X >> lv;
        Y = s->Y >> lv;
        Z = s->Z >> lv;
    };
    int X, Y, Z;
    if (keep_cat || !VC_CONV_FUSE01)
    {
        if ((rc = conv(0, in)) != VC_OK || (rc = conv(1, s->act[0])) != VC_OK)
            return rc;
    }
    else
    {
        // inference: layers 0 and 1 in one kernel (layer 0's activations are
        // not stored)
        const int64_t nb = (int64_t)((s->X + kCBX - 1) / kCBX) * ((s->Y + kCBY - 1) / kCBY) *
                           ((s->Z + 15) / 16);
        if (!s->inf_occ)
            return vc_fail(ctx, VC_ERR_INVALID, "inference without its volume");
        conv01_lds<kCBX, kCBY><<<conv_grid(nb), kBlock, 0, S>>>(
            s->inf_occ, s->inf_edt, s->inf_n[0], s->inf_n[1], s->inf_n[2], s->act[1],
            reinterpret_cast<const bf16x8*>(s->W4),
            s->theta + param_offset(0) + 27LL * kLayers[0].cin * kLayers[0].cout,
            reinterpret_cast<const bf16x8*>(s->Wf[1]),
            s->theta + param_offset(1) + 27LL * kLayers[1].cin * kLayers[1].cout, s->X, s->Y,
            s->Z, level_box(s, 0), conv_off32(s->Y, s->Z, 16));
        VC_CHECK_LAUNCH(ctx);
    }
    // encoder
    bf16* pools[3] = {s->p1, s->p2, s->p3};
    const int pc[3] = {16, 32, 48};
    for (int lv = 0; lv < 3; ++lv)
    {
        dims(lv, X, Y, Z);
        const int64_t n = level_voxels(s, lv + 1) * pc[lv];
        pool_fwd<<<grid_for(n / 8), kBlock, 0, S>>>(s->act[2 * lv + 1], pools[lv],
                                                       pc[lv], X, Y, Z);
        VC_CHECK_LAUNCH(ctx);
        if ((rc = conv(2 * lv + 2, pools[lv])) != VC_OK ||
            (rc = conv(2 * lv + 3, s->act[2 * lv + 2])) != VC_OK)
            return rc;
    }
    // decoder: level 2, 1, 0
    bf16* cats[3] = {s->cat2, s->cat1, s->cat0};
    const int c1s[3] = {64, 48, 32}, c2s[3] = {48, 32, 16};
    const bf16* ups[3] = {s->act[7], s->act[9], s->act[11]};
    const bf16* skips[3] = {s->act[5], s->act[3], s->act[1]};
    for (int d = 0; d < 3; ++d)
    {
        const int lv = 2 - d;
        dims(lv, X, Y, Z);
        if (keep_cat)
        {
            const int64_t n = level_voxels(s, lv) * (c1s[d] + c2s[d]);
            concat_fwd<<<grid_for(n), kBlock, 0, S>>>(ups[d], c1s[d], skips[d],
                                                         c2s[d], cats[d], X, Y, Z);
            VC_CHECK_LAUNCH(ctx);
            rc = conv(8 + 2 * d, cats[d]);
        }
        else
        {
            const int l = 8 + 2 * d;
            const LayerSpec& L = kLayers[l];
            rc = launch_conv(ctx, ups[d], L.cin, s->act[l], L.cout, s->Wf[l],
                             s->theta + param_offset(l) + 27LL * L.cin * L.cout, X, Y,
                             Z, s->kc[l], 1, level_box(s, L.level), skips[d], c1s[d]);
        }
        if (rc != VC_OK)
            return rc;
        if (d == 2 && !keep_cat)
        {
            // inference: the last layer with the head fused (its activations
            // are not kept)
            const LayerSpec& L = kLayers[13];
            const int64_t nb = (int64_t)((X + kCBX - 1) / kCBX) * ((Y + kCBY - 1) / kCBY) *
                               ((Z + 16 * kCNZT - 1) / (16 * kCNZT));
            conv_fwd_lds<1, kCBX, kCBY, kCNZT, kCSC, false, true, true>
                <<<conv_grid(nb), kBlock, 0, S>>>(
                    s->act[12], nullptr, 0, cinp_of(L.cin), nullptr,
                    reinterpret_cast<const bf16x8*>(s->Wf[13]),
                    s->theta + param_offset(13) + 27LL * L.cin * L.cout, X, Y, Z,
                    level_box(s, 0), conv_off32(Y, Z, cinp_of(L.cin)),
                    s->theta + param_offset(14), s->out, s->inf_core);
            VC_CHECK_LAUNCH(ctx);
            return VC_OK;
        }
        if ((rc = conv(9 + 2 * d, s->act[8 + 2 * d])) != VC_OK)
            return rc;
    }
    const int64_t V0 = level_voxels(s, 0);
    head_fwd<<<blocks_for(V0), kBlock, 0, S>>>(s->act[13],
                                                s->theta + param_offset(14),
                                                s->out, V0);
    VC_CHECK_LAUNCH(ctx);
    return VC_OK;
}

// layer l backward: gy (gradient of its ReLU output, in place) -> masked,
// weight gradient, and (unless gx is null) the input gradient into gx
int layer_back(vc_ctx* ctx, vc_surrogate* s, int l, bf16* gy, const bf16* x,
               bf16* gx)
{
    const LayerSpec& L = kLayers[l];
    const int X = s->X >> L.level, Y = s->Y >> L.level, Z = s->Z >> L.level;
    const int64_t V = (int64_t)X * Y * Z;
    hipStream_t S = ctx->stream;
    relu_mask<<<grid_for(V * L.cout), kBlock, 0, S>>>(gy, s->act[l],
                                                         V * L.cout);
    VC_CHECK_LAUNCH(ctx);
    int rc = layer_wgrad(ctx, s, l, x, gy);
    if (rc != VC_OK || !gx)
        return rc;
    return launch_conv(ctx, gy, L.cout, gx, cinp_of(L.cin), s->Wd[l], nullptr,
                       X, Y, Z, s->kcd[l], 0, level_box(s, L.level));
}

// inside voxels, one add per workgroup (a capped grid strides over them)
__global__ void __launch_bounds__(kBlock)
    count_occ(const uint8_t* __restrict__ occ, int64_t N,
              unsigned long long* __restrict__ cnt)
{
    __shared__ unsigned long long sh[kBlock / 64];
    unsigned long long c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * blockDim.x)
        c += occ[i] != 0;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        c += __shfl_xor(c, d);
    if ((threadIdx.x & 63) == 0)
        sh[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned long long t = 0;
        for (int w = 0; w < kBlock / 64; ++w)
            t += sh[w];
        if (t)
            atomicAdd(cnt, t);
    }
}

// dst = sum of part[0..n) (fixed order) / max(count, 1)
__global__ void __launch_bounds__(kBlock)
    loss_reduce(const float* __restrict__ part, int n,
                const unsigned long long* __restrict__ cnt,
                float* __restrict__ dst)
{
    __shared__ double red[kBlock];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += kBlock)
        s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = kBlock / 2; k > 0; k >>= 1)
    {
        if (threadIdx.x < k)
            red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        const unsigned long long c = *cnt;
        *dst = (float)(red[0] / (double)(c ? c : 1ull));
    }
}

// backward of the loss on a region equal to the (padded) volume; adds the
// gradient into s->grad and the loss into *loss_dev
int backward(vc_ctx* ctx, vc_surrogate* s, const vc_vol* v, float* loss_dev,
             unsigned long long* cnt)
{
    hipStream_t S = ctx->stream;
    const int64_t V0 = level_voxels(s, 0);
    const unsigned nb = blocks_for(V0);
    VC_HIP(ctx, hipMemsetAsync(cnt, 0, 8, S));
    count_occ<<<(unsigned)std::min<int64_t>(blocks_for(v->N), 4096), kBlock, 0, S>>>(
        v->occ, v->N, cnt);
    VC_CHECK_LAUNCH(ctx);
    loss_bwd<<<nb, kBlock, 0, S>>>(s->out, v->occ, v->lam2, v->nx, v->ny,
                                    v->nz, s->X, s->Y, s->Z, cnt, s->dout,
                                    s->part);
    VC_CHECK_LAUNCH(ctx);
    loss_reduce<<<1, kBlock, 0, S>>>(s->part, (int)nb, cnt, loss_dev);
    VC_CHECK_LAUNCH(ctx);
    constexpr int NP = kHeadIn * kHeadOut + kHeadOut;
    head_bwd<<<nb, kBlock, 0, S>>>(s->act[13], s->dout,
                                    s->theta + param_offset(14), s->gact[13],
                                    V0, s->part);
    VC_CHECK_LAUNCH(ctx);
    colsum_reduce<<<1, 64, 0, S>>>(s->part, (int)nb, NP,
                                    s->grad + param_offset(14));
    VC_CHECK_LAUNCH(ctx);
    int rc;
    auto dims = [&](int lv, int& X, int& Y, int& Z) {
        X = s->X >> lv;
        Y = s->Y >> lv;
        Z = s->Z >> lv;
    };
    int X, Y, Z;
    // decoder levels 0, 1, 2: (second conv, first conv, concat input, its
    // gradient, up channels, coarse gradient target, skip gradient target)
    struct Dec
    {
        int lb, la, lv, C, C1;
        bf16 *cat, *gcat, *gup, *gskip;
    };
    const Dec dec[3] = {
        {13, 12, 0, 48, 32, s->cat0, s->gcat0, s->gact[11], s->gact[1]},
        {11, 10, 1, 80, 48, s->cat1, s->gcat1, s->gact[9], s->gact[3]},
        {9, 8, 2, 112, 64, s->cat2, s->gcat2, s->gact[7], s->gact[5]}};
    for (const Dec& d : dec)
    {
        if ((rc = layer_back(ctx, s, d.lb, s->gact[d.lb], s->act[d.la],
                             s->gact[d.la])) != VC_OK)
            return rc;
        if ((rc = layer_back(ctx, s, d.la, s->gact[d.la], d.cat, d.gcat)) !=
            VC_OK)
            return rc;
        dims(d.lv, X, Y, Z);
        const int64_t nu = level_voxels(s, d.lv + 1) * d.C1;
        split_up<<<grid_for(nu), kBlock, 0, S>>>(d.gcat, d.C, d.C1, d.gup, X,
                                                    Y, Z);
        VC_CHECK_LAUNCH(ctx);
        const int64_t Vl = level_voxels(s, d.lv);
        split_skip<<<grid_for(Vl * (d.C - d.C1)), kBlock, 0, S>>>(
            d.gcat, d.C, d.C1, d.gskip, Vl);
        VC_CHECK_LAUNCH(ctx);
    }
    // encoder levels 3, 2, 1: the pooled input's gradient adds into the
    // gradient of the level above's second conv output
    bf16* pools[3] = {s->p1, s->p2, s->p3};
    bf16* gpools[3] = {s->gp1, s->gp2, s->gp3};
    const int pc[3] = {16, 32, 48};
    for (int lv = 3; lv >= 1; --lv)
    {
        const int lb = 2 * lv + 1, la = 2 * lv;
        if ((rc = layer_back(ctx, s, lb, s->gact[lb], s->act[la],
                             s->gact[la])) != VC_OK)
            return rc;
        if ((rc = layer_back(ctx, s, la, s->gact[la], pools[lv - 1],
                             gpools[lv - 1])) != VC_OK)
            return rc;
        dims(lv - 1, X, Y, Z);
        const int64_t n = level_voxels(s, lv - 1) * pc[lv - 1];
        pool_bwd_acc<<<grid_for(n), kBlock, 0, S>>>(
            s->gact[2 * lv - 1], gpools[lv - 1], pc[lv - 1], X, Y, Z);
        VC_CHECK_LAUNCH(ctx);
    }
    if ((rc = layer_back(ctx, s, 1, s->gact[1], s->act[0], s->gact[0])) !=
        VC_OK)
        return rc;
    return layer_back(ctx, s, 0, s->gact[0], s->x8, nullptr);
}

int round8(int n) { return (n + 7) / 8 * 8; }

// forward + backward of one volume (gradient added into s->grad)
int train_volume(vc_ctx* ctx, vc_surrogate* s, const vc_vol* v, float* loss_dev,
                 unsigned long long* cnt)
{
    int rc = ensure_ws(ctx, s, round8(v->nx), round8(v->ny), round8(v->nz),
                       true);
    if (rc != VC_OK)
        return rc;
    s->org[0] = s->org[1] = s->org[2] = 0;
    s->pad[0] = s->X;
    s->pad[1] = s->Y;
    s->pad[2] = s->Z;
    hipStream_t S = ctx->stream;
    prep_input<<<blocks_for(level_voxels(s, 0)), kBlock, 0, S>>>(
        v->occ, v->edt, v->nx, v->ny, v->nz, 0, 0, 0, s->x0, s->x8, s->X, s->Y, s->Z);
    VC_CHECK_LAUNCH(ctx);
    if ((rc = forward(ctx, s, true)) != VC_OK)
        return rc;
    return backward(ctx, s, v, loss_dev, cnt);
}

int check_labelled(vc_ctx* ctx, const vc_vol* v)
{
    if (!v->edt_done || !v->lam2 || v->measure_kind == 0)
        return vc_fail(ctx, VC_ERR_STATE,
                       "surrogate training needs vc_edt and vc_measure");
    return VC_OK;
}

int ensure_vol_out(vc_ctx* ctx, vc_surrogate* s, int64_t N)
{
    if (s->vN == N)
        return VC_OK;
    VC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    vc_free(s->vlam);
    vc_free(s->vrel);
    vc_free(s->vmeas);
    s->vlam = s->vrel = s->vmeas = nullptr;
    s->vN = 0;
    int rc = vc_alloc(ctx, (void**)&s->vlam, 4 * (size_t)N);
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&s->vrel, 4 * (size_t)N);
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&s->vmeas, 4 * (size_t)N);
    if (rc == VC_OK)
        s->vN = N;
    return rc;
}

// network outputs of a whole volume into s->vlam / s->vrel
int run_infer(vc_ctx* ctx, vc_surrogate* s, const vc_vol* v, int tile_core)
{
    int rc = ensure_vol_out(ctx, s, v->N);
    if (rc != VC_OK)
        return rc;
    hipStream_t S = ctx->stream;
    const int dims[3] = {v->nx, v->ny, v->nz};
    int amax = dims[0] > dims[1] ? dims[0] : dims[1];
    amax = amax > dims[2] ? amax : dims[2];
    // a conv launch covers its region with 2 threads per voxel, counted in
    // 32 bits: a whole volume beyond that runs as tiles of 512 (tiled and
    // whole-volume passes give the same outputs bit for bit)
    const int64_t whole_vox =
        (int64_t)round8(dims[0]) * round8(dims[1]) * round8(dims[2]);
    const bool fits = 2 * whole_vox < (int64_t(1) << 32);
    const bool whole = (tile_core <= 0 || tile_core >= amax) && fits;
    const int T = whole ? 0
                        : round8(tile_core > 0 && tile_core < amax ? tile_core : 512);
    int R[3];
    for (int a = 0; a < 3; ++a)
        R[a] = whole ? round8(dims[a]) : T + 2 * kHalo;
    if ((rc = ensure_ws(ctx, s, R[0], R[1], R[2], false)) != VC_OK)
        return rc;
    const int step[3] = {whole ? dims[0] : T, whole ? dims[1] : T,
                         whole ? dims[2] : T};
    const int halo = whole ? 0 : kHalo;
    const int64_t VR = level_voxels(s, 0);
    for (int tx = 0; tx < dims[0]; tx += step[0])
        for (int ty = 0; ty < dims[1]; ty += step[1])
            for (int tz = 0; tz < dims[2]; tz += step[2])
            {
                const int ox = tx - halo, oy = ty - halo, oz = tz - halo;
                s->org[0] = ox;
                s->org[1] = oy;
                s->org[2] = oz;
                for (int a = 0; a < 3; ++a)
                    s->pad[a] = round8(dims[a]);
                if (VC_CONV_FUSE01)
                {
                    // layers 0 and 1 read the volume directly
                    s->inf_occ = v->occ;
                    s->inf_edt = v->edt;
                    s->inf_n[0] = v->nx;
                    s->inf_n[1] = v->ny;
                    s->inf_n[2] = v->nz;
                }
                else
                {
                    prep_input<<<blocks_for(VR), kBlock, 0, S>>>(
                        v->occ, v->edt, v->nx, v->ny, v->nz, ox, oy, oz, s->x0,
                        nullptr, s->X, s->Y, s->Z);
                    VC_CHECK_LAUNCH(ctx);
                }
                // the fused head writes the core's lam / rel itself
                s->inf_core = CoreOut{s->vlam, s->vrel, v->nx, v->ny, v->nz, tx, ty, tz,
                                      tx + step[0], ty + step[1], tz + step[2]};
                rc = forward(ctx, s, false);
                s->inf_occ = nullptr;
                s->inf_edt = nullptr;
                s->inf_core = CoreOut{};
                if (rc != VC_OK)
                    return rc;
                if (!VC_CONV_FUSED_HEAD_CORE)
                {
                    copy_core<<<blocks_for(VR), kBlock, 0, S>>>(
                        s->out, s->X, s->Y, s->Z, ox, oy, oz, tx, ty, tz,
                        tx + step[0], ty + step[1], tz + step[2], s->vlam, s->vrel,
                        v->nx, v->ny, v->nz);
                    VC_CHECK_LAUNCH(ctx);
                }
            }
    return VC_OK;
}

} // namespace

extern "C" {

int vc_surrogate_create(vc_ctx* ctx, uint64_t seed, vc_surrogate** out)
{
    if (!ctx || !out)
        return VC_ERR_INVALID;
    *out = nullptr;
    VC_HIP(ctx, hipSetDevice(ctx->device));
    vc_surrogate* s = new (std::nothrow) vc_surrogate();
    if (!s)
        return VC_ERR_NOMEM;
    s->ctx = ctx;
    s->np = total_params();
    int rc = VC_OK;
    if (s->np != VC_SURROGATE_PARAMS)
        rc = vc_fail(ctx, VC_ERR_INVALID, "parameter count %lld",
                     (long long)s->np);
    const size_t pb = 4 * (size_t)s->np;
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&s->theta, pb);
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&s->grad, pb);
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&s->m1, pb);
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&s->m2, pb);
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&s->ctr, 64 * 8);
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&s->lossbuf, 4 * 64);
    for (int l = 0; l < 14 && rc == VC_OK; ++l)
    {
        const LayerSpec& L = kLayers[l];
        const int cinp = cinp_of(L.cin);
        s->kc[l] = round4(27 * cinp / 8);
        rc = vc_alloc(ctx, (void**)&s->Wf[l], 2 * (size_t)s->kc[l] * L.cout * 8);
        if (rc == VC_OK && l > 0)
        {
            s->kcd[l] = round4(27 * L.cout / 8);
            rc = vc_alloc(ctx, (void**)&s->Wd[l],
                          2 * (size_t)s->kcd[l] * cinp * 8);
        }
    }
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&s->W4, 2 * (size_t)8 * 16 * 8);
    if (rc == VC_OK)
    {
        // He-uniform initial weights (oracle/unet_ref.py init_params)
        std::vector<float> th((size_t)s->np, 0.0f);
        int64_t o = 0;
        for (int l = 0; l <= 14; ++l)
        {
            const bool head = l == 14;
            const int ci = head ? kHeadIn : kLayers[l].cin;
            const int co = head ? kHeadOut : kLayers[l].cout;
            const int taps = head ? 1 : 27;
            const int64_t n = (int64_t)taps * ci * co;
            float a = sqrtf(6.0f / (float)(taps * ci));
            if (head)
                a = a * 0.1f;
            for (int64_t j = o; j < o + n; ++j)
            {
                const uint64_t r =
                    splitmix64(((seed & 0xFFFFFFFFull) << 32) + (uint64_t)j);
                const float u = (float)(r >> 40) * (1.0f / 16777216.0f);
                th[(size_t)j] = (2.0f * u - 1.0f) * a;
            }
            o += n + co;
        }
        hipError_t e = hipMemcpy(s->theta, th.data(), pb, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemset(s->m1, 0, pb);
        if (e == hipSuccess)
            e = hipMemset(s->m2, 0, pb);
        if (e != hipSuccess)
            rc = vc_fail(ctx, VC_ERR_HIP, "surrogate init: %s",
                         hipGetErrorString(e));
    }
    if (rc == VC_OK)
        rc = pack_weights(ctx, s);
    if (rc != VC_OK)
    {
        vc_surrogate_destroy(ctx, s);
        return rc;
    }
    *out = s;
    return VC_OK;
}

int vc_surrogate_destroy(vc_ctx* ctx, vc_surrogate* s)
{
    if (!ctx || !s)
        return VC_ERR_INVALID;
    (void)hipStreamSynchronize(ctx->stream);
    free_ws(s);
    vc_free(s->theta);
    vc_free(s->grad);
    vc_free(s->m1);
    vc_free(s->m2);
    vc_free(s->ctr);
    vc_free(s->lossbuf);
    vc_free(s->vlam);
    vc_free(s->vrel);
    vc_free(s->vmeas);
    for (int l = 0; l < 14; ++l)
    {
        vc_free(s->Wf[l]);
        vc_free(s->Wd[l]);
    }
    vc_free(s->W4);
    delete s;
    return VC_OK;
}

int vc_surrogate_get_params(vc_ctx* ctx, vc_surrogate* s, float* host,
                            int64_t n)
{
    if (!ctx || !s || !host || n != s->np)
        return VC_ERR_INVALID;
    VC_HIP(ctx, hipMemcpyAsync(host, s->theta, 4 * (size_t)n,
                               hipMemcpyDeviceToHost, ctx->stream));
    VC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VC_OK;
}

int vc_surrogate_set_params(vc_ctx* ctx, vc_surrogate* s, const float* host,
                            int64_t n)
{
    if (!ctx || !s || !host || n != s->np)
        return VC_ERR_INVALID;
    VC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    VC_HIP(ctx, hipMemcpy(s->theta, host, 4 * (size_t)n,
                          hipMemcpyHostToDevice));
    VC_HIP(ctx, hipMemset(s->m1, 0, 4 * (size_t)n));
    VC_HIP(ctx, hipMemset(s->m2, 0, 4 * (size_t)n));
    s->step = 0;
    return pack_weights(ctx, s);
}

int vc_surrogate_save(vc_ctx* ctx, vc_surrogate* s, const char* path)
{
    if (!ctx || !s || !path)
        return VC_ERR_INVALID;
    std::vector<float> th((size_t)s->np);
    int rc = vc_surrogate_get_params(ctx, s, th.data(), s->np);
    if (rc != VC_OK)
        return rc;
    FILE* f = std::fopen(path, "wb");
    if (!f)
        return vc_fail(ctx, VC_ERR_INVALID, "cannot write %s", path);
    const int64_t n = s->np;
    bool ok = std::fwrite("VCUNET01", 1, 8, f) == 8 &&
              std::fwrite(&n, 8, 1, f) == 1 &&
              std::fwrite(th.data(), 4, th.size(), f) == th.size();
    ok = (std::fclose(f) == 0) && ok;
    return ok ? VC_OK : vc_fail(ctx, VC_ERR_INVALID, "short write to %s", path);
}

int vc_surrogate_load(vc_ctx* ctx, vc_surrogate* s, const char* path)
{
    if (!ctx || !s || !path)
        return VC_ERR_INVALID;
    FILE* f = std::fopen(path, "rb");
    if (!f)
        return vc_fail(ctx, VC_ERR_INVALID, "cannot read %s", path);
    char magic[8];
    int64_t n = 0;
    std::vector<float> th((size_t)s->np);
    bool ok = std::fread(magic, 1, 8, f) == 8 &&
              std::memcmp(magic, "VCUNET01", 8) == 0 &&
              std::fread(&n, 8, 1, f) == 1 && n == s->np &&
              std::fread(th.data(), 4, th.size(), f) == th.size();
    std::fclose(f);
    if (!ok)
        return vc_fail(ctx, VC_ERR_INVALID, "%s is not a weights file of this "
                                            "network", path);
    return vc_surrogate_set_params(ctx, s, th.data(), s->np);
}

int vc_surrogate_gradients(vc_ctx* ctx, vc_surrogate* s, vc_vol* vol,
                           float* grads, float* loss)
{
    if (!ctx || !s || !vol || vol->ctx != ctx)
        return VC_ERR_INVALID;
    int rc = check_labelled(ctx, vol);
    if (rc != VC_OK)
        return rc;
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    fill_f32<<<grid_for(s->np), kBlock, 0, S>>>(s->grad, s->np, 0.0f);
    VC_CHECK_LAUNCH(ctx);
    if ((rc = train_volume(ctx, s, vol, s->lossbuf, s->ctr)) != VC_OK)
        return rc;
    if (grads)
        VC_HIP(ctx, hipMemcpyAsync(grads, s->grad, 4 * (size_t)s->np,
                                   hipMemcpyDeviceToHost, S));
    float l = 0.0f;
    VC_HIP(ctx, hipMemcpyAsync(&l, s->lossbuf, 4, hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    if (loss)
        *loss = l;
    return VC_OK;
}

int vc_surrogate_train_step(vc_ctx* ctx, vc_surrogate* s, vc_vol* const* vols,
                            int n, float lr, float* loss)
{
    if (!ctx || !s || !vols || n < 1 || n > 64 || !(lr > 0.0f))
        return VC_ERR_INVALID;
    for (int i = 0; i < n; ++i)
    {
        if (!vols[i] || vols[i]->ctx != ctx)
            return VC_ERR_INVALID;
        int rc = check_labelled(ctx, vols[i]);
        if (rc != VC_OK)
            return rc;
    }
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    fill_f32<<<grid_for(s->np), kBlock, 0, S>>>(s->grad, s->np, 0.0f);
    VC_CHECK_LAUNCH(ctx);
    for (int i = 0; i < n; ++i)
    {
        int rc = train_volume(ctx, s, vols[i], s->lossbuf + i, s->ctr + i);
        if (rc != VC_OK)
            return rc;
    }
    s->step++;
    const float bc1 = 1.0f - powf(0.9f, (float)s->step);
    const float bc2 = 1.0f - powf(0.999f, (float)s->step);
    adam_step<<<blocks_for(s->np), kBlock, 0, S>>>(s->theta, s->grad, s->m1,
                                                    s->m2, s->np, lr,
                                                    1.0f / (float)n, bc1, bc2);
    VC_CHECK_LAUNCH(ctx);
    int rc = pack_weights(ctx, s);
    if (rc != VC_OK)
        return rc;
    if (loss)
    {
        float l[64];
        VC_HIP(ctx, hipMemcpyAsync(l, s->lossbuf, 4 * (size_t)n,
                                   hipMemcpyDeviceToHost, S));
        VC_HIP(ctx, hipStreamSynchronize(S));
        double t = 0.0;
        for (int i = 0; i < n; ++i)
            t += l[i];
        *loss = (float)(t / n);
    }
    return VC_OK;
}

int vc_surrogate_infer(vc_ctx* ctx, vc_surrogate* s, vc_vol* vol,
                       int tile_core, float* lambda_out, float* relerr_out)
{
    if (!ctx || !s || !vol || vol->ctx != ctx)
        return VC_ERR_INVALID;
    if (!vol->edt_done)
        return vc_fail(ctx, VC_ERR_STATE, "surrogate inference needs vc_edt");
    VC_HIP(ctx, hipSetDevice(ctx->device));
    int rc = run_infer(ctx, s, vol, tile_core);
    if (rc != VC_OK)
        return rc;
    hipStream_t S = ctx->stream;
    if (lambda_out)
        VC_HIP(ctx, hipMemcpyAsync(lambda_out, s->vlam, 4 * (size_t)vol->N,
                                   hipMemcpyDeviceToHost, S));
    if (relerr_out)
        VC_HIP(ctx, hipMemcpyAsync(relerr_out, s->vrel, 4 * (size_t)vol->N,
                                   hipMemcpyDeviceToHost, S));
    VC_HIP(ctx, hipStreamSynchronize(S));
    return VC_OK;
}

int vc_surrogate_measure(vc_ctx* ctx, vc_surrogate* s, vc_vol* vol,
                         float max_rel_err, int tile_core,
                         vc_surrogate_stats* stats)
{
    if (!ctx || !s || !vol || vol->ctx != ctx || !(max_rel_err >= 0.0f))
        return VC_ERR_INVALID;
    if (!vol->edt_done)
        return vc_fail(ctx, VC_ERR_STATE, "surrogate measure needs vc_edt");
    VC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t S = ctx->stream;
    hipEvent_t ev[4];
    for (int k = 0; k < 4; ++k)
        VC_HIP(ctx, hipEventCreate(&ev[k]));
    int rc = VC_OK;
    hipError_t e = hipEventRecord(ev[0], S);
    if (e == hipSuccess && (rc = run_infer(ctx, s, vol, tile_core)) == VC_OK)
    {
        e = hipEventRecord(ev[1], S);
        if (e == hipSuccess)
            e = hipMemsetAsync(s->ctr, 0, 3 * 8, S);
        if (e == hipSuccess)
        {
            gate_count<<<(unsigned)std::min<int64_t>(blocks_for(vol->N), 4096), kBlock, 0, S>>>(
                vol->occ, s->vrel, max_rel_err, vol->N, s->ctr);
            e = hipGetLastError();
        }
        unsigned long long hc[3] = {0, 0, 0};
        if (e == hipSuccess)
            e = hipMemcpyAsync(hc, s->ctr, 3 * 8, hipMemcpyDeviceToHost, S);
        if (e == hipSuccess)
            e = hipStreamSynchronize(S);
        if (e == hipSuccess)
            e = hipEventRecord(ev[2], S);
        if (e == hipSuccess && hc[1] > 0 && vol->measure_kind != 1)
        {
            // the exact measure for the fallback voxels
            rc = vc_measure(ctx, vol);
        }
        if (e == hipSuccess && rc == VC_OK)
            e = hipEventRecord(ev[3], S);
        if (e == hipSuccess && rc == VC_OK)
        {
            gate_mix<<<blocks_for(vol->N), kBlock, 0, S>>>(
                vol->occ, s->vlam, s->vrel, hc[1] > 0 ? vol->lam2 : nullptr,
                max_rel_err, vol->N, s->vmeas);
            e = hipGetLastError();
        }
        if (e == hipSuccess && rc == VC_OK)
            rc = vc_set_measure_f32(ctx, vol, s->vmeas);
        hipEvent_t e_end;
        if (e == hipSuccess && rc == VC_OK && stats &&
            (e = hipEventCreate(&e_end)) == hipSuccess)
        {
            e = hipEventRecord(e_end, S);
            if (e == hipSuccess)
                e = hipEventSynchronize(e_end);
            float t01 = 0, t23 = 0, t3e = 0;
            if (e == hipSuccess)
                e = hipEventElapsedTime(&t01, ev[0], ev[1]);
            if (e == hipSuccess)
                e = hipEventElapsedTime(&t23, ev[2], ev[3]);
            if (e == hipSuccess)
                e = hipEventElapsedTime(&t3e, ev[3], e_end);
            (void)hipEventDestroy(e_end);
            stats->inside = (int64_t)hc[2];
            stats->accepted = (int64_t)hc[0];
            stats->fallback = (int64_t)hc[1];
            stats->ms_infer = t01;
            stats->ms_exact = t23;
            stats->ms_mix = t3e;
        }
    }
    for (int k = 0; k < 4; ++k)
        (void)hipEventDestroy(ev[k]);
    if (rc != VC_OK)
        return rc;
    if (e != hipSuccess)
        return vc_fail(ctx, VC_ERR_HIP, "surrogate measure: %s",
                       hipGetErrorString(e));
    return VC_OK;
}

int vc_surrogate_activation(vc_ctx* ctx, vc_surrogate* s, int which,
                            float* host, int64_t n)
{
    if (!ctx || !s || !host || which < 0 || which > 14 || s->X == 0)
        return VC_ERR_INVALID;
    if ((which == 0 || which == 13) && !s->full_acts)
        return vc_fail(ctx, VC_ERR_INVALID,
                       "activation %d is not stored by an inference (layers 0 and 1 run "
                       "in one kernel; the head is evaluated inside layer 13's "
                       "convolution)", which);
    const int64_t V =
        which == 14 ? level_voxels(s, 0) : level_voxels(s, kLayers[which].level);
    const int C = which == 14 ? 2 : kLayers[which].cout;
    if (n != V * C)
        return vc_fail(ctx, VC_ERR_INVALID, "activation %d has %lld values",
                       which, (long long)(V * C));
    hipStream_t S = ctx->stream;
    if (which == 14)
    {
        VC_HIP(ctx, hipMemcpyAsync(host, s->out, 4 * (size_t)n,
                                   hipMemcpyDeviceToHost, S));
    }
    else
    {
        float* tmp = nullptr;
        int rc = vc_alloc(ctx, (void**)&tmp, 4 * (size_t)n);
        if (rc != VC_OK)
            return rc;
        bf16_to_f32<<<grid_for(n), kBlock, 0, S>>>(s->act[which], tmp, n);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(host, tmp, 4 * (size_t)n, hipMemcpyDeviceToHost,
                               S);
        if (e == hipSuccess)
            e = hipStreamSynchronize(S);
        vc_free(tmp);
        if (e != hipSuccess)
            return vc_fail(ctx, VC_ERR_HIP, "activation: %s",
                           hipGetErrorString(e));
    }
    VC_HIP(ctx, hipStreamSynchronize(S));
    return VC_OK;
}

} // extern "C"
