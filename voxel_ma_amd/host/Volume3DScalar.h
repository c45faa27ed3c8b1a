// Volume3DScalar.h / DenseVolume -- the reference's volume input interface
// (include/Volume3DScalar.h:11-34, include/densevolume.h:13-48), with a dense
// in-memory implementation.  Occupancy = getDataAt > 0, out of bounds = 0
// (include/spaceinfo.h:122-135); occupancy() hands the whole grid to the
// device in one copy instead of one virtual call per voxel.
#pragma once

#include <memory>
#include <vector>

#include "commondefs.h"

class Volume3DScalar
{
public:
    Volume3DScalar() : m_sizes(0) {}
    virtual ~Volume3DScalar() = default;
    virtual double getDataAt(int x, int y, int z) const = 0;
    int getSizeX() const { return m_sizes[0]; }
    int getSizeY() const { return m_sizes[1]; }
    int getSizeZ() const { return m_sizes[2]; }
    const xform& getVoxToModelMat() const { return m_vox_to_mesh; }
    const xform& getModeltoVoxMat() const { return m_mesh_to_vox; }
    // uint8 occupancy, C order [x][y][z] (z fastest)
    virtual std::vector<uint8_t> occupancy() const
    {
        std::vector<uint8_t> o((size_t)getSizeX() * getSizeY() * getSizeZ());
        size_t i = 0;
        for (int x = 0; x < getSizeX(); ++x)
            for (int y = 0; y < getSizeY(); ++y)
                for (int z = 0; z < getSizeZ(); ++z)
                    o[i++] = getDataAt(x, y, z) > 0.0 ? 1 : 0;
        return o;
    }

protected:
    ivec3 m_sizes;
    xform m_vox_to_mesh, m_mesh_to_vox;
};

class DenseVolume : public Volume3DScalar
{
public:
    DenseVolume() = default;
    DenseVolume(int nx, int ny, int nz, std::vector<float> values)
        : m_data(std::move(values))
    {
        m_sizes = ivec3(nx, ny, nz);
        if (m_data.size() != (size_t)nx * ny * nz)
            m_data.assign((size_t)nx * ny * nz, 0.0f);
    }
    static std::shared_ptr<DenseVolume> fromOccupancy(int nx, int ny, int nz,
                                                      const uint8_t* occ)
    {
        std::vector<float> v((size_t)nx * ny * nz);
        for (size_t i = 0; i < v.size(); ++i)
            v[i] = occ[i] ? 1.0f : 0.0f;
        return std::make_shared<DenseVolume>(nx, ny, nz, std::move(v));
    }
    double getDataAt(int x, int y, int z) const override
    {
        return m_data[((size_t)x * getSizeY() + y) * getSizeZ() + z];
    }
    std::vector<uint8_t> occupancy() const override
    {
        std::vector<uint8_t> o(m_data.size());
        for (size_t i = 0; i < o.size(); ++i)
            o[i] = m_data[i] > 0.0f ? 1 : 0;
        return o;
    }

private:
    std::vector<float> m_data;
};
