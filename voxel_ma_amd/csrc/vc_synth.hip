// vc_synth.hip -- seeded synthetic input volumes on the device (benchmark
// plumbing, not part of the thinning path).
//
// Blurred-noise volume: N(0,1) noise from a counter-based generator
// (splitmix64 -> Box-Muller), separable Gaussian blur (odd tap count,
// periodic boundary, taps applied in ascending order), then
// inside iff value > the exact k-th smallest value (radix select on the
// order-preserving integer image of the float bits), then a 1-voxel empty
// shell.  Deterministic for a given seed on a given device.
#include <cmath>
#include <cstring>
#include <vector>

#include "vc_internal.h"
#include "../../include/vc_synth.h"

namespace
{

constexpr int kBlock = 256;
constexpr int kMaxTaps = 255;

struct Taps
{
    float w[kMaxTaps];
};

__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void __launch_bounds__(kBlock)
    noise_kernel(float* out, int64_t N, uint64_t seed)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const uint64_t r = splitmix64(seed * 0x632BE59BD9B4E019ull + (uint64_t)i);
    // two 24-bit uniforms in (0, 1]
    const float u1 = ((float)(r >> 40) + 1.0f) * (1.0f / 16777216.0f);
    const float u2 = (float)((r >> 16) & 0xFFFFFFull) * (1.0f / 16777216.0f);
    out[i] = sqrtf(-2.0f * logf(u1)) * cosf(6.28318530717958647692f * u2);
}

// blur along one axis: element (a, b, c) of a [A][B][C] view with the blur
// axis being `axis` (0 = A, 1 = B, 2 = C); periodic boundary
__global__ void __launch_bounds__(kBlock)
    blur_kernel(const float* __restrict__ in, float* __restrict__ out, int nx,
                int ny, int nz, int axis, Taps taps, int ntaps)
{
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int x = blockIdx.z;
    if (z >= nz)
        return;
    const int rad = ntaps / 2;
    const int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
    const int64_t stride = axis == 0 ? (int64_t)ny * nz : (axis == 1 ? nz : 1);
    const int pos = axis == 0 ? x : (axis == 1 ? y : z);
    const int64_t base = ((int64_t)x * ny + y) * nz + z - (int64_t)pos * stride;
    float acc = 0.0f;
    for (int t = 0; t < ntaps; ++t)
    {
        int p = pos + t - rad;
        p %= n;
        if (p < 0)
            p += n;
        acc = fmaf(taps.w[t], in[base + (int64_t)p * stride], acc);
    }
    out[((int64_t)x * ny + y) * nz + z] = acc;
}

__device__ __forceinline__ uint32_t float_key(float f)
{
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// histogram of the digit (key >> shift) & mask among keys whose bits above
// `hishift` equal `prefix`
__global__ void __launch_bounds__(kBlock)
    radix_hist(const float* __restrict__ v, int64_t N, uint32_t prefix,
               int hishift, int shift, uint32_t mask, uint32_t* hist)
{
    extern __shared__ uint32_t h[];
    for (uint32_t b = threadIdx.x; b <= mask; b += blockDim.x)
        h[b] = 0;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += stride)
    {
        const uint32_t k = float_key(v[i]);
        const bool match =
            hishift >= 32 ? true : ((k >> hishift) == prefix);
        if (match)
            atomicAdd(&h[(k >> shift) & mask], 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b <= mask; b += blockDim.x)
        if (h[b])
            atomicAdd(&hist[b], h[b]);
}

__global__ void __launch_bounds__(kBlock)
    threshold_kernel(const float* __restrict__ v, uint8_t* __restrict__ occ,
                     int nx, int ny, int nz, uint32_t thr_key, int shell)
{
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int x = blockIdx.z;
    if (z >= nz)
        return;
    const int64_t i = ((int64_t)x * ny + y) * nz + z;
    bool in = float_key(v[i]) > thr_key;
    if (shell && (x == 0 || y == 0 || z == 0 || x == nx - 1 || y == ny - 1 ||
                  z == nz - 1))
        in = false;
    occ[i] = in ? 1 : 0;
}

// Union of solid primitives tested at voxel centres.  Each record is
// VC_SOLID_WORDS doubles (see vc_synth.h); every operation is a separately
// rounded IEEE double operation in the stated order (no contraction), so a
// host restatement in the same order gives the same bytes.
__global__ void __launch_bounds__(kBlock)
    solids_kernel(const double* __restrict__ prims, int n_prims, uint8_t* occ,
                  int nx, int ny, int nz)
{
#pragma clang fp contract(off)
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int x = blockIdx.z;
    if (z >= nz)
        return;
    uint8_t in = 0;
    const bool shell = x == 0 || y == 0 || z == 0 || x == nx - 1 ||
                       y == ny - 1 || z == nz - 1;
    for (int p = 0; p < n_prims && !in && !shell; ++p)
    {
        const double* r = prims + (size_t)p * VC_SOLID_WORDS;
        const double dx = (double)x - r[1], dy = (double)y - r[2],
                     dz = (double)z - r[3];
        double q[3];
        for (int j = 0; j < 3; ++j)
            q[j] = (dx * r[4 + j] + dy * r[7 + j]) + dz * r[10 + j];
        if (r[0] == 0.0)
        {
            const double t0 = q[0] / r[13], t1 = q[1] / r[14], t2 = q[2] / r[15];
            in = ((t0 * t0 + t1 * t1) + t2 * t2) <= 1.0;
        }
        else if (r[0] == 2.0)
        {
            const double rho = sqrt(q[0] * q[0] + q[1] * q[1]) - r[13];
            in = (rho * rho + q[2] * q[2]) <= r[14] * r[14];
        }
        else
        {
            const double h = r[14];
            const double t = q[0] < -h ? -h : (q[0] > h ? h : q[0]);
            const double e = q[0] - t;
            in = ((e * e + q[1] * q[1]) + q[2] * q[2]) <= r[13] * r[13];
        }
    }
    occ[((int64_t)x * ny + y) * nz + z] = in;
}

} // namespace

extern "C" int vc_synth_solids(vc_ctx* ctx, uint8_t* d_occ, int nx, int ny,
                               int nz, const double* prims, int n_prims)
{
    if (!ctx || !d_occ || nx <= 0 || ny <= 0 || nz <= 0 || n_prims < 0 ||
        (n_prims > 0 && !prims))
        return VC_ERR_INVALID;
    if (ny > 65535 || nx > 65535)
        return vc_fail(ctx, VC_ERR_INVALID, "vc_synth_solids: axis too long");
    VC_HIP(ctx, hipSetDevice(ctx->device));
    double* dp = nullptr;
    const size_t bytes = sizeof(double) * VC_SOLID_WORDS * (size_t)n_prims;
    int rc = vc_alloc(ctx, (void**)&dp, bytes);
    if (rc != VC_OK)
        return rc;
    hipError_t e = hipSuccess;
    if (n_prims > 0)
        e = hipMemcpyAsync(dp, prims, bytes, hipMemcpyHostToDevice,
                           ctx->stream);
    if (e == hipSuccess)
    {
        const dim3 g((unsigned)((nz + kBlock - 1) / kBlock), (unsigned)ny,
                     (unsigned)nx);
        solids_kernel<<<g, kBlock, 0, ctx->stream>>>(dp, n_prims, d_occ, nx,
                                                     ny, nz);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    vc_free(dp);
    if (e != hipSuccess)
        return vc_fail(ctx, VC_ERR_HIP, "vc_synth_solids: %s",
                       hipGetErrorString(e));
    return VC_OK;
}

extern "C" int vc_synth_blurred_noise(vc_ctx* ctx, uint8_t* d_occ, int nx,
                                      int ny, int nz, float sigma,
                                      float fill, uint64_t seed, int shell,
                                      float* threshold_out)
{
    if (!ctx || !d_occ || nx <= 0 || ny <= 0 || nz <= 0 || !(sigma > 0.0f) ||
        !(fill > 0.0f && fill < 1.0f))
        return vc_fail(ctx, VC_ERR_INVALID, "vc_synth_blurred_noise: bad args");
    VC_HIP(ctx, hipSetDevice(ctx->device));
    const int rad = (int)(4.0f * sigma + 0.5f);
    const int ntaps = 2 * rad + 1;
    if (ntaps > kMaxTaps)
        return vc_fail(ctx, VC_ERR_INVALID, "sigma too large");
    Taps taps{};
    {
        double sum = 0.0;
        std::vector<double> w(ntaps);
        for (int t = 0; t < ntaps; ++t)
        {
            double xx = (double)(t - rad) / sigma;
            w[t] = std::exp(-0.5 * xx * xx);
            sum += w[t];
        }
        for (int t = 0; t < ntaps; ++t)
            taps.w[t] = (float)(w[t] / sum);
    }
    const int64_t N = (int64_t)nx * ny * nz;
    float *a = nullptr, *b = nullptr;
    uint32_t* hist = nullptr;
    int rc = vc_alloc(ctx, (void**)&a, sizeof(float) * (size_t)N);
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&b, sizeof(float) * (size_t)N);
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&hist, sizeof(uint32_t) * 4096);
    hipStream_t S = ctx->stream;
    hipError_t e = hipSuccess;
    uint32_t thr_key = 0;
    if (rc == VC_OK)
    {
        noise_kernel<<<(unsigned)((N + kBlock - 1) / kBlock), kBlock, 0, S>>>(
            a, N, seed);
        e = hipGetLastError();
        dim3 g((unsigned)((nz + kBlock - 1) / kBlock), (unsigned)ny,
               (unsigned)nx);
        for (int axis = 2; axis >= 0 && e == hipSuccess; --axis)
        {
            blur_kernel<<<g, kBlock, 0, S>>>(a, b, nx, ny, nz, axis, taps,
                                              ntaps);
            e = hipGetLastError();
            std::swap(a, b);
        }
        // exact k-th smallest (1-based) with k = ceil((1 - fill) * N)
        int64_t k = (int64_t)std::ceil((1.0 - (double)fill) * (double)N);
        if (k < 1)
            k = 1;
        const int digits[3] = {11, 11, 10};
        int shift = 32;
        uint32_t prefix = 0;
        std::vector<uint32_t> hh(4096);
        for (int d = 0; d < 3 && e == hipSuccess; ++d)
        {
            const int hishift = shift;
            shift -= digits[d];
            const uint32_t mask = (1u << digits[d]) - 1u;
            e = hipMemsetAsync(hist, 0, sizeof(uint32_t) * 4096, S);
            if (e != hipSuccess)
                break;
            radix_hist<<<1024, kBlock, sizeof(uint32_t) * (mask + 1), S>>>(
                a, N, prefix, hishift, shift, mask, hist);
            e = hipGetLastError();
            if (e == hipSuccess)
                e = hipMemcpyAsync(hh.data(), hist, sizeof(uint32_t) * (mask + 1),
                                   hipMemcpyDeviceToHost, S);
            if (e == hipSuccess)
                e = hipStreamSynchronize(S);
            if (e != hipSuccess)
                break;
            uint32_t dsel = 0;
            for (uint32_t bb = 0; bb <= mask; ++bb)
            {
                if ((int64_t)hh[bb] >= k)
                {
                    dsel = bb;
                    break;
                }
                k -= hh[bb];
            }
            prefix = (prefix << digits[d]) | dsel;
        }
        thr_key = prefix;
        if (e == hipSuccess)
        {
            threshold_kernel<<<g, kBlock, 0, S>>>(a, d_occ, nx, ny, nz, thr_key,
                                                   shell);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipStreamSynchronize(S);
    }
    vc_free(a);
    vc_free(b);
    vc_free(hist);
    if (rc != VC_OK)
        return rc;
    if (e != hipSuccess)
        return vc_fail(ctx, VC_ERR_HIP, "synth: %s", hipGetErrorString(e));
    if (threshold_out)
    {
        uint32_t u = (thr_key & 0x80000000u) ? (thr_key & 0x7FFFFFFFu) : ~thr_key;
        float f;
        std::memcpy(&f, &u, 4);
        *threshold_out = f;
    }
    return VC_OK;
}
