// kifi/ICP.h -- rigid alignment of a raw depth frame to a rendered point buffer
// (reference: KinectFission/kifi/ICP.h:13-43, ICP.cpp:9-265): four window-search
// Horn steps on the GPU.  Poses are eye->world, as ICP.cpp uses them.  The
// projective point-to-plane tracker of the fused pipeline is kifi::Pipeline's.
#pragma once

#include <cstdint>
#include <vector>

#include "kifi/DepthSensorParams.h"
#include "kifi/Integrator.h"
#include "kifi/util/math.h"
#include "kifi/util/vector2d.h"

namespace kifi {

class ICP {
public:
    util::float4x4 Align(const util::vector2d<float> &rawDepthMap, const util::float4x4 &rawDepthMapEyeToWorldGuess,
                         const util::vector2d<util::float3> &synthDepthBuffer,
                         const util::float4x4 &synthDepthBufferEyeToWorld, const DepthSensorParams &cameraParams)
    {
        const kifi_intrinsics c = detail::frame_camera(cameraParams, rawDepthMap.width(), rawDepthMap.height());
        util::float4x4 out;
        detail::check(kifi_icp_align_points_f32(detail::context(), rawDepthMap.data(), &c,
                                                rawDepthMapEyeToWorldGuess.data(),
                                            &synthDepthBuffer.data()->x, (int)synthDepthBuffer.width(),
                                            (int)synthDepthBuffer.height(), synthDepthBufferEyeToWorld.data(), 4,
                                            out.data(), &m_stats),
                      "kifi_icp_align_points");
        return out;
    }

    const kifi_icp_stats &LastStats() const { return m_stats; }

private:
    kifi_icp_stats m_stats{};
};

} // namespace kifi
