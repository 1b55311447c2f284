// kifi/Pipeline.h -- the frame loop (reference: KinectFission/kifi/Pipeline.h:15-58,
// Pipeline.cpp:7-68).  Integrate(frame, worldToEye) is the reference's active
// branch (given poses, Pipeline.cpp:32-35); Integrate(frame) is its tracking
// branch (Pipeline.cpp:36-55), run on the device as projective point-to-plane
// ICP against a model rendered from the volume.
#pragma once

#include <cstdint>
#include <vector>

#include "kifi/DepthSensorParams.h"
#include "kifi/Integrator.h"
#include "kifi/Mesher.h"
#include "kifi/Volume.h"
#include "kifi/util/math.h"
#include "kifi/util/vector2d.h"

namespace kifi {

class Pipeline {
public:
    Pipeline(const DepthSensorParams &cameraParams, int volumeResolution = 512, float volumeSideLength = 4.0f,
             float truncationMargin = 0.02f, int keyBits = 32, int storage = KIFI_STORAGE_F32,
             const kifi_icp_params *icp = nullptr)
        : m_camParams(cameraParams), m_cam(cameraParams.Intrinsics()), m_eyeToWorld(util::float4x4::identity())
    {
        kifi_volume_desc d;
        d.res = volumeResolution;
        d.side = volumeSideLength;
        d.trunc = truncationMargin;
        d.key_bits = keyBits;
        d.storage = storage;
        detail::check(kifi_pipeline_create(detail::context(), &d, &m_cam, icp, &m_h), "kifi_pipeline_create");
        m_volume = new kifi::Volume(kifi::Volume::Borrow(kifi_pipeline_volume(m_h), d));
        m_eyeToWorld.col3.z = 2.0f;                 // Pipeline.cpp:19-21
        SetPose(m_eyeToWorld);
    }
    ~Pipeline()
    {
        delete m_volume;
        kifi_pipeline_destroy(m_h);
    }
    Pipeline(const Pipeline &) = delete;
    Pipeline &operator=(const Pipeline &) = delete;

    void Integrate(const util::vector2d<float> &rawDepthMap, const util::float4x4 &worldToEye)
    {
        detail::frame_camera(m_camParams, rawDepthMap.width(), rawDepthMap.height());
        detail::check(kifi_pipeline_integrate_f32(m_h, rawDepthMap.data(), 0, worldToEye.data()),
                      "kifi_pipeline_integrate");
        m_eyeToWorld = worldToEye;
        util::invert_transform(m_eyeToWorld);
    }

    // tracking: the pose comes from ICP against the previous frame's model
    void Integrate(const util::vector2d<float> &rawDepthMap)
    {
        detail::frame_camera(m_camParams, rawDepthMap.width(), rawDepthMap.height());
        detail::check(kifi_pipeline_integrate_f32(m_h, rawDepthMap.data(), 0, nullptr), "kifi_pipeline_integrate");
        m_eyeToWorld = Pose();
    }

    void Mesh(std::vector<util::float3> &outVertices) { m_mesher.Mesh(*m_volume, outVertices); }
    void Mesh(std::vector<util::float3> &outVertices, std::vector<unsigned> &outIndices)
    {
        m_mesher.Mesh(*m_volume, outVertices, outIndices);
    }

    const kifi::Volume &Volume() const { return *m_volume; }
    const util::float4x4 &EyeToWorld() const { return m_eyeToWorld; }

    // eye->world of the last frame, from the device (synchronous)
    util::float4x4 Pose() const
    {
        double p[16];
        detail::check(kifi_pipeline_pose(m_h, p), "kifi_pipeline_pose");
        util::float4x4 r;
        for (int i = 0; i < 16; ++i)
            r.data()[i] = (float)p[i];
        return r;
    }
    void SetPose(const util::float4x4 &eyeToWorld)
    {
        double p[16];
        for (int i = 0; i < 16; ++i)
            p[i] = eyeToWorld.data()[i];
        detail::check(kifi_pipeline_set_pose(m_h, p), "kifi_pipeline_set_pose");
    }
    kifi_icp_stats Stats() const
    {
        kifi_icp_stats s;
        detail::check(kifi_pipeline_stats(m_h, &s), "kifi_pipeline_stats");
        return s;
    }

private:
    DepthSensorParams m_camParams;
    kifi_intrinsics m_cam;
    kifi_pipeline *m_h = nullptr;
    kifi::Volume *m_volume = nullptr;
    Mesher m_mesher;
    util::float4x4 m_eyeToWorld;
};

} // namespace kifi
