// vc_internal.h -- shared host/device declarations of libvoxelcore (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <string>

#include "../../include/vc_api.h"

// ---------------------------------------------------------------- context
struct vc_ctx
{
    int device = 0;
    int cu_count = 0;
    hipStream_t stream = nullptr;
    uint32_t* d_lut = nullptr; // simple-point table, 2^21 words
    std::string err;
};

// packed voxel coordinate: x << (by+bz) | y << bz | z, fits in 31 bits
struct vc_pack
{
    int bx, by, bz;
    uint32_t my, mz;
};

struct vc_vol
{
    vc_ctx* ctx = nullptr;
    int nx = 0, ny = 0, nz = 0;
    int64_t N = 0;
    vc_pack pk{};
    uint8_t* occ = nullptr;   // N, 0/1
    int32_t* edt = nullptr;   // N
    uint32_t* ft = nullptr;   // N, packed coordinate, 0xFFFFFFFF = none
    uint32_t* lam2 = nullptr; // N
    float* meas_f = nullptr;  // N, caller-supplied measure
    int measure_kind = 0;     // 0 none, 1 exact lam2, 2 float
    uint8_t* lvl = nullptr;   // N
    uint8_t* label = nullptr; // N
    // thinning state: bit planes packed along z (64 voxels per word)
    uint64_t* alive = nullptr;  // alive voxels
    uint64_t* elig = nullptr;   // eligible at the current level (lvl <= k)
    uint8_t* dflag = nullptr;   // 2 x words: revisit flags per z parity
    uint32_t* sweep_any = nullptr; // per sweep of a level: removed anything
    uint32_t* hcnt = nullptr;   // pinned host staging for counters
    uint64_t* tstats = nullptr; // device stats accumulators
    // boundary stage results (computed once per volume)
    bool bnd_done = false;
    float* bnd_sites = nullptr;  // 3 floats per site
    int32_t* bnd_quads = nullptr; // 4 site indices per quad
    int64_t bnd_nsites = 0, bnd_nquads = 0;
    vc_euler bnd_euler{};
    bool edt_done = false;
    bool thin_done = false;
};

static constexpr uint32_t VC_NONE = 0xFFFFFFFFu;

// ---------------------------------------------------------------- errors
int vc_fail(vc_ctx* ctx, int code, const char* fmt, ...);

#define VC_HIP(ctx, call)                                                       \
    do                                                                          \
    {                                                                           \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess)                                                   \
            return vc_fail((ctx), VC_ERR_HIP, "%s: %s (%s:%d)", #call,          \
                           hipGetErrorString(e_), __FILE__, __LINE__);          \
    } while (0)

#define VC_CHECK_LAUNCH(ctx) VC_HIP(ctx, hipGetLastError())

int vc_alloc(vc_ctx* ctx, void** p, size_t bytes);
void vc_free(void* p);

// ---------------------------------------------------------------- stages
int vc_lut_build(vc_ctx* ctx);
int vc_edt_run(vc_ctx* ctx, vc_vol* vol);
int vc_measure_run(vc_ctx* ctx, vc_vol* vol);
int vc_thin_run(vc_ctx* ctx, vc_vol* vol, const float* levels, int n,
                vc_thin_stats* st);
int vc_boundary_run(vc_ctx* ctx, vc_vol* vol, float* sites, int64_t site_cap,
                    int64_t* n_sites, int32_t* quads, int64_t quad_cap,
                    int64_t* n_quads, vc_euler* eu);
int vc_ft_to_linear(vc_ctx* ctx, vc_vol* vol, int32_t* d_dst);

// host helpers
int64_t vc_level_threshold(float level);

__host__ __device__ inline uint32_t vc_pack_xyz(const vc_pack& p, int x, int y,
                                                int z)
{
    return ((uint32_t)x << (p.by + p.bz)) | ((uint32_t)y << p.bz) | (uint32_t)z;
}
__host__ __device__ inline void vc_unpack_xyz(const vc_pack& p, uint32_t v,
                                              int& x, int& y, int& z)
{
    z = (int)(v & p.mz);
    y = (int)((v >> p.bz) & p.my);
    x = (int)(v >> (p.by + p.bz));
}
