// vc_octree.hip -- sparse octree input (include/vc_octree.h), gfx950.
//
// The node stream is parsed on the host into flat arrays (type, sign byte,
// index of the 8-entry child block, node corner and edge); the dense
// (R+1)^3 occupancy is then built on the device, one thread per voxel
// walking the tree from the root (depth log2 R).
#include <cstring>
#include <new>
#include <vector>

#include "vc_internal.h"
#include "../../include/vc_octree.h"

struct vc_octree
{
    int32_t R = 0;
    std::vector<uint8_t> type, val;
    std::vector<int32_t> child; // first of 8 children per node, -1 if none
    std::vector<int32_t> kids;  // child blocks, (i, j, k) with k fastest
    std::vector<int32_t> ox, oy, oz, len;
    vc_octree_info info{};
};

namespace
{

constexpr int kBlock = 256;
constexpr uint8_t kInternal = 0, kEmpty = 1, kLeaf = 2;

struct Reader
{
    const uint8_t* p;
    size_t n, at = 0;
    bool get(void* dst, size_t k)
    {
        if (at + k > n)
            return false;
        std::memcpy(dst, p + at, k);
        at += k;
        return true;
    }
};

// pre-order parse; false on a truncated stream, a bad type or a depth that
// exceeds log2 R
bool parse_node(Reader& rd, vc_octree* t, int ox, int oy, int oz, int len,
                bool sog, int depth)
{
    uint8_t type = 0;
    if (!rd.get(&type, 1) || depth > 31)
        return false;
    const int32_t id = (int32_t)t->type.size();
    t->type.push_back(type);
    t->val.push_back(0);
    t->child.push_back(-1);
    t->ox.push_back(ox);
    t->oy.push_back(oy);
    t->oz.push_back(oz);
    t->len.push_back(len);
    if (type == kLeaf || type == kEmpty)
    {
        uint8_t v = 0;
        if (!rd.get(&v, 1))
            return false;
        t->val[(size_t)id] = v;
        if (type == kLeaf)
        {
            t->info.n_leaf++;
            if (sog)
            {
                float dump[3];
                if (!rd.get(dump, sizeof(dump)))
                    return false;
            }
        }
        else
            t->info.n_empty++;
        return true;
    }
    if (type != kInternal || len < 2)
        return false;
    t->info.n_internal++;
    const int32_t base = (int32_t)t->kids.size();
    t->child[(size_t)id] = base;
    t->kids.resize(t->kids.size() + 8, -1);
    const int h = len / 2;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k)
            {
                t->kids[(size_t)base + i * 4 + j * 2 + k] = (int32_t)t->type.size();
                if (!parse_node(rd, t, ox + i * h, oy + j * h, oz + k * h, h,
                                sog, depth + 1))
                    return false;
            }
    return true;
}

// sign byte of the unit cell (x, y, z), 0 <= x, y, z < R: the leaf's byte,
// 0xFF / 0 for an empty node (sign 1 / other), 0 when not found
__host__ __device__ inline uint8_t cell_value(const uint8_t* type,
                                              const uint8_t* val,
                                              const int32_t* child,
                                              const int32_t* kids, int R,
                                              int x, int y, int z)
{
    if (x < 0 || y < 0 || z < 0 || x >= R || y >= R || z >= R)
        return 0;
    int32_t n = 0;
    int ox = 0, oy = 0, oz = 0, len = R;
    while (type[n] == kInternal && len > 1)
    {
        const int h = len >> 1;
        const int i = x >= ox + h, j = y >= oy + h, k = z >= oz + h;
        ox += i * h;
        oy += j * h;
        oz += k * h;
        len = h;
        n = kids[child[n] + i * 4 + j * 2 + k];
    }
    if (type[n] == kLeaf)
        return val[n];
    if (type[n] == kEmpty)
        return val[n] == 1 ? (uint8_t)0xFF : (uint8_t)0;
    return 0;
}

// OctreeVolume::getDataAt: a coordinate equal to R reads the cell below
// with the corner bit x -> 4, y -> 2, z -> 1; inside iff that bit is 0
__host__ __device__ inline bool octree_inside(const uint8_t* type,
                                              const uint8_t* val,
                                              const int32_t* child,
                                              const int32_t* kids, int R,
                                              int x, int y, int z)
{
    int off = 0;
    if (x == R)
    {
        x = R - 1;
        off |= 4;
    }
    if (y == R)
    {
        y = R - 1;
        off |= 2;
    }
    if (z == R)
    {
        z = R - 1;
        off |= 1;
    }
    const uint8_t v = cell_value(type, val, child, kids, R, x, y, z);
    return ((v >> off) & 1) == 0;
}

__global__ void __launch_bounds__(kBlock)
    octree_occupancy(const uint8_t* __restrict__ type,
                     const uint8_t* __restrict__ val,
                     const int32_t* __restrict__ child,
                     const int32_t* __restrict__ kids, int R,
                     uint8_t* __restrict__ occ)
{
    const int n = R + 1;
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, x = blockIdx.z;
    if (z >= n)
        return;
    occ[((int64_t)x * n + y) * n + z] =
        octree_inside(type, val, child, kids, R, x, y, z) ? 1 : 0;
}

} // namespace

extern "C" {

int vc_octree_parse(const uint8_t* bytes, size_t n, int format,
                    vc_octree** out, vc_octree_info* info)
{
    if (!bytes || !out || (format != VC_OCTREE_SOF && format != VC_OCTREE_SOG))
        return VC_ERR_INVALID;
    *out = nullptr;
    vc_octree* t = new (std::nothrow) vc_octree();
    if (!t)
        return VC_ERR_NOMEM;
    Reader rd{bytes, n};
    bool ok = true;
    if (format == VC_OCTREE_SOG)
    {
        static const char title[] = "SOG.Format 1.0";
        uint8_t hdr[128];
        ok = rd.get(hdr, sizeof(hdr)) &&
             std::memcmp(hdr, title, sizeof(title)) == 0;
        if (ok)
        {
            std::memcpy(t->info.lower_left, hdr + sizeof(title), 12);
            std::memcpy(&t->info.bbox_len, hdr + sizeof(title) + 12, 4);
        }
    }
    ok = ok && rd.get(&t->R, 4) && t->R >= 1 && t->R <= (1 << 10) &&
         (t->R & (t->R - 1)) == 0;
    if (ok)
    {
        t->info.resolution = t->R;
        ok = parse_node(rd, t, 0, 0, 0, t->R, format == VC_OCTREE_SOG, 0);
    }
    if (!ok)
    {
        delete t;
        return VC_ERR_INVALID;
    }
    if (info)
        *info = t->info;
    *out = t;
    return VC_OK;
}

int vc_octree_free(vc_octree* t)
{
    delete t;
    return VC_OK;
}

int vc_octree_value(const vc_octree* t, int x, int y, int z, double* value)
{
    if (!t || !value || x < 0 || y < 0 || z < 0 || x > t->R || y > t->R ||
        z > t->R)
        return VC_ERR_INVALID;
    *value = octree_inside(t->type.data(), t->val.data(), t->child.data(),
                           t->kids.data(), t->R, x, y, z)
                 ? 1.0
                 : -1.0;
    return VC_OK;
}

int vc_volume_from_octree(vc_ctx* ctx, const vc_octree* t, vc_vol** out)
{
    if (!ctx || !t || !out)
        return VC_ERR_INVALID;
    *out = nullptr;
    VC_HIP(ctx, hipSetDevice(ctx->device));
    const size_t nn = t->type.size(), nk = t->kids.size();
    uint8_t *dt = nullptr, *dv = nullptr, *docc = nullptr;
    int32_t *dc = nullptr, *dk = nullptr;
    const int n = t->R + 1;
    int rc = vc_alloc(ctx, (void**)&dt, nn);
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&dv, nn);
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&dc, 4 * nn);
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&dk, 4 * nk + 4);
    if (rc == VC_OK)
        rc = vc_alloc(ctx, (void**)&docc, (size_t)n * n * n);
    hipStream_t S = ctx->stream;
    hipError_t e = hipSuccess;
    if (rc == VC_OK)
    {
        e = hipMemcpyAsync(dt, t->type.data(), nn, hipMemcpyHostToDevice, S);
        if (e == hipSuccess)
            e = hipMemcpyAsync(dv, t->val.data(), nn, hipMemcpyHostToDevice, S);
        if (e == hipSuccess)
            e = hipMemcpyAsync(dc, t->child.data(), 4 * nn,
                               hipMemcpyHostToDevice, S);
        if (e == hipSuccess && nk)
            e = hipMemcpyAsync(dk, t->kids.data(), 4 * nk,
                               hipMemcpyHostToDevice, S);
        if (e == hipSuccess)
        {
            const dim3 g((unsigned)((n + kBlock - 1) / kBlock), (unsigned)n,
                         (unsigned)n);
            octree_occupancy<<<g, kBlock, 0, S>>>(dt, dv, dc, dk, t->R, docc);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            rc = vc_volume_from_device(ctx, docc, n, n, n, out);
        if (e == hipSuccess && rc == VC_OK)
            e = hipStreamSynchronize(S);
    }
    vc_free(dt);
    vc_free(dv);
    vc_free(dc);
    vc_free(dk);
    vc_free(docc);
    if (rc != VC_OK)
        return rc;
    if (e != hipSuccess)
    {
        if (*out)
            vc_volume_release(ctx, *out);
        *out = nullptr;
        return vc_fail(ctx, VC_ERR_HIP, "octree volume: %s",
                       hipGetErrorString(e));
    }
    return VC_OK;
}

int vc_octree_leaf_centers(const vc_octree* t, float* xyz, int64_t cap,
                           int64_t* n_out)
{
    if (!t || !n_out)
        return VC_ERR_INVALID;
    int64_t k = 0;
    for (size_t i = 0; i < t->type.size(); ++i)
    {
        if (t->type[i] != kLeaf || t->val[i] == 0 || t->val[i] == 0xFF)
            continue;
        if (xyz && k < cap)
        {
            xyz[3 * k] = (float)t->ox[i] + 0.5f;
            xyz[3 * k + 1] = (float)t->oy[i] + 0.5f;
            xyz[3 * k + 2] = (float)t->oz[i] + 0.5f;
        }
        ++k;
    }
    *n_out = k;
    return (xyz && k > cap) ? VC_ERR_CAPACITY : VC_OK;
}

} // extern "C"
