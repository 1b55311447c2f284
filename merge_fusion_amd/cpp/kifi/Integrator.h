// kifi/Integrator.h -- depth frame -> sparse TSDF volume (reference:
// KinectFission/kifi/Integrator.h:15-37, Integrator.cpp:26-71).  A frame in
// metres is used as it is; a uint16 millimetre frame is converted on the device
// with the reference's own (float)mm * 0.001f (DepthStream.cpp:89-90).
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#include "kifi/DepthSensorParams.h"
#include "kifi/Volume.h"
#include "kifi/util/math.h"
#include "kifi/util/vector2d.h"

namespace kifi {

namespace detail {
inline kifi_intrinsics frame_camera(const DepthSensorParams &p, std::size_t w, std::size_t h)
{
    kifi_intrinsics c = p.Intrinsics();
    if ((std::size_t)c.width != w || (std::size_t)c.height != h) {
        std::fprintf(stderr, "kifi: frame is %zux%zu, camera is %dx%d\n", w, h, c.width, c.height);
        std::abort();
    }
    return c;
}
} // namespace detail

class Integrator {
public:
    void Integrate(Volume &volume, const util::vector2d<float> &frame, const DepthSensorParams &cameraParams,
                   const util::float4x4 &worldToEye)
    {
        const kifi_intrinsics c = detail::frame_camera(cameraParams, frame.width(), frame.height());
        detail::check(kifi_integrate_f32(volume.Handle(), frame.data(), &c, worldToEye.data()), "kifi_integrate");
    }

    // millimetre frames, as the sensor delivers them
    void Integrate(Volume &volume, const util::vector2d<uint16_t> &frameMM, const DepthSensorParams &cameraParams,
                   const util::float4x4 &worldToEye)
    {
        const kifi_intrinsics c = detail::frame_camera(cameraParams, frameMM.width(), frameMM.height());
        detail::check(kifi_integrate(volume.Handle(), frameMM.data(), &c, worldToEye.data()), "kifi_integrate");
    }
};

} // namespace kifi
