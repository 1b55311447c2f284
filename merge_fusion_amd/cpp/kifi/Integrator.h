// kifi/Integrator.h -- depth frame -> sparse TSDF volume (reference:
// KinectFission/kifi/Integrator.h:15-37, Integrator.cpp:26-71).  The GPU path
// reads depth as uint16 millimetres and converts with (float)mm * 0.001f, the
// reference's own conversion (DepthStream.cpp:89-90); a float frame in metres is
// rounded to millimetres on the host first, which reproduces every frame that
// conversion can produce exactly (all 65536 values round-trip).
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
inline void metres_to_mm(const util::vector2d<float> &m, std::vector<uint16_t> &mm)
{
    mm.resize(m.size());
    for (std::size_t i = 0; i < m.size(); ++i) {
        const float v = m[i] * 1000.0f;
        mm[i] = (v > 0.0f && v < 65535.5f) ? (uint16_t)std::lrintf(v) : (uint16_t)0;   // NaN -> 0 (invalid)
    }
}

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
        detail::metres_to_mm(frame, m_mm);
        const kifi_intrinsics c = detail::frame_camera(cameraParams, frame.width(), frame.height());
        detail::check(kifi_integrate(volume.Handle(), m_mm.data(), &c, worldToEye.data()), "kifi_integrate");
    }

    // millimetre frames, as the sensor delivers them
    void Integrate(Volume &volume, const util::vector2d<uint16_t> &frameMM, const DepthSensorParams &cameraParams,
                   const util::float4x4 &worldToEye)
    {
        const kifi_intrinsics c = detail::frame_camera(cameraParams, frameMM.width(), frameMM.height());
        detail::check(kifi_integrate(volume.Handle(), frameMM.data(), &c, worldToEye.data()), "kifi_integrate");
    }

private:
    std::vector<uint16_t> m_mm;
};

} // namespace kifi
