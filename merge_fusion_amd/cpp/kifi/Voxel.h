// kifi/Voxel.h -- a voxel's running TSDF sum and weight (reference: KinectFission/kifi/Voxel.h:11-57).
// With KIFI_STORAGE_F16U8 volumes the library stores the distance in binary16
// and the weight in 8 bits; Volume::Data() hands both layouts back as this type.
#pragma once

#include <cassert>

namespace kifi {

struct Voxel {
    Voxel() : m_distance(0.0f), m_weight(0.0f) {}
    Voxel(float sum, float weight) : m_distance(sum), m_weight(weight) {}

    float Distance() const
    {
        assert(m_weight > 0.0f);
        return m_distance / m_weight;
    }
    float Weight() const { return m_weight; }
    void Update(float distance)
    {
        m_distance += distance;
        m_weight++;
    }

private:
    float m_distance;   // sum of the truncated distances
    float m_weight;
};

} // namespace kifi
