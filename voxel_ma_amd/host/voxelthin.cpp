#include "voxelthin.h"

#include <stdexcept>

#include <hip/hip_runtime.h>

#include "vc_device.h"

void VoxelThinning::setup(const std::shared_ptr<Volume3DScalar>& vol)
{
    reset();
    if (!vol)
        throw std::invalid_argument("VoxelThinning::setup: null volume");
    const std::vector<uint8_t> occ = vol->occupancy();
    m_n[0] = vol->getSizeX();
    m_n[1] = vol->getSizeY();
    m_n[2] = vol->getSizeZ();
    vcdev::check(vc_volume_upload(vcdev::context(), occ.data(), m_n[0], m_n[1],
                                  m_n[2], &m_vol),
                 "vc_volume_upload");
}

void VoxelThinning::reset()
{
    if (m_vol)
        vc_volume_release(vcdev::context(), m_vol);
    m_vol = nullptr;
    m_has_measure = false;
    m_labels.clear();
}

void VoxelThinning::computeMeasure()
{
    vc_ctx* c = vcdev::context();
    vcdev::check(vc_edt(c, m_vol, 0), "vc_edt");
    vcdev::check(vc_measure(c, m_vol), "vc_measure");
    m_has_measure = true;
}

void VoxelThinning::assignMeasure(const std::vector<float>& m)
{
    const size_t N = (size_t)m_n[0] * m_n[1] * m_n[2];
    if (m.size() != N)
        throw std::invalid_argument("VoxelThinning::assignMeasure: size");
    // stage through the device: vc_set_measure_f32 takes device memory
    vc_ctx* c = vcdev::context();
    float* d = nullptr;
    if (hipMalloc(&d, N * sizeof(float)) != hipSuccess)
        throw std::runtime_error("hipMalloc for the measure failed");
    int rc = hipMemcpy(d, m.data(), N * sizeof(float), hipMemcpyHostToDevice) ==
                     hipSuccess
                 ? vc_set_measure_f32(c, m_vol, d)
                 : VC_ERR_HIP;
    vc_ctx_synchronize(c);
    (void)hipFree(d);
    vcdev::check(rc, "vc_set_measure_f32");
    m_has_measure = true;
}

void VoxelThinning::prune(const std::vector<float>& levels)
{
    if (!m_has_measure)
        computeMeasure();
    vc_ctx* c = vcdev::context();
    vcdev::check(vc_thin_levels(c, m_vol, levels.data(), (int)levels.size(),
                                &m_stats),
                 "vc_thin_levels");
    m_labels.resize((size_t)m_n[0] * m_n[1] * m_n[2]);
    vcdev::check(vc_get(c, m_vol, VC_BUF_LABEL, m_labels.data(),
                        (int64_t)m_labels.size()),
                 "vc_get(labels)");
}

void VoxelThinning::getRemainedVoxels(int k, std::vector<ivec3>& vox) const
{
    vox.clear();
    size_t i = 0;
    for (int x = 0; x < m_n[0]; ++x)
        for (int y = 0; y < m_n[1]; ++y)
            for (int z = 0; z < m_n[2]; ++z, ++i)
                if (m_labels[i] == 1 || m_labels[i] > 2 + k)
                    vox.push_back(ivec3(x, y, z));
}

std::vector<int32_t> VoxelThinning::squaredDistances() const
{
    std::vector<int32_t> d((size_t)m_n[0] * m_n[1] * m_n[2]);
    vcdev::check(vc_get(vcdev::context(), m_vol, VC_BUF_EDT, d.data(),
                        (int64_t)(4 * d.size())),
                 "vc_get(edt)");
    return d;
}

std::vector<uint32_t> VoxelThinning::lambda2() const
{
    std::vector<uint32_t> d((size_t)m_n[0] * m_n[1] * m_n[2]);
    vcdev::check(vc_get(vcdev::context(), m_vol, VC_BUF_LAM2, d.data(),
                        (int64_t)(4 * d.size())),
                 "vc_get(lambda2)");
    return d;
}
