// kifi/Renderer.h -- z-buffered point splat (reference: KinectFission/kifi/Renderer.h:13-33,
// Renderer.cpp:39-80).  The output buffer keeps the size the caller gave it.
#pragma once

#include <vector>

#include "kifi/detail.h"
#include "kifi/util/math.h"
#include "kifi/util/vector2d.h"

namespace kifi {

class Renderer {
public:
    void Bin(const std::vector<util::float3> &pointCloud, const util::float4x4 &worldToClip,
             util::vector2d<util::float3> &outPointBuffer)
    {
        detail::check(kifi_bin(detail::context(), pointCloud.empty() ? nullptr : &pointCloud.data()->x,
                               pointCloud.size(), worldToClip.data(), (int)outPointBuffer.width(),
                               (int)outPointBuffer.height(), &outPointBuffer.data()->x),
                      "kifi_bin");
    }
};

} // namespace kifi
