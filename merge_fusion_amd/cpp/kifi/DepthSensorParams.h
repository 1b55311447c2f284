// kifi/DepthSensorParams.h -- camera model (reference: KinectFission/kifi/DepthSensorParams.h:23-45,
// DepthSensorParams.cpp:12-80).
#pragma once

#include "kifi/util/math.h"
#include "kifi.h"

namespace kifi {

enum KinectDepthSensorResolution { KinectDepthSensorResolution320x240, KinectDepthSensorResolution640x480 };
enum KinectDepthSensorMode { KinectDepthSensorModeNear, KinectDepthSensorModeFar };

class DepthSensorParams {
public:
    DepthSensorParams(util::int2 resolutionPixels, util::float2 focalLengthPixels,
                      util::float2 principalPointPixels, util::float2 sensibleDistanceRangeMeters)
        : m_res(resolutionPixels), m_fl(focalLengthPixels), m_pp(principalPointPixels),
          m_range(sensibleDistanceRangeMeters) {}

    util::int2 ResolutionPixels() const { return m_res; }
    util::float2 FocalLengthPixels() const { return m_fl; }
    util::float2 PrincipalPointPixels() const { return m_pp; }
    util::float2 SensibleRangeMeters() const { return m_range; }

    // OpenGL-style projection from right-handed eye space (DepthSensorParams.cpp:47-63)
    util::float4x4 EyeToClipRH() const
    {
        util::float4x4 r(0.0f);
        const float W = (float)m_res.x, H = (float)m_res.y, zn = m_range.x, zf = m_range.y;
        r(0, 0) = 2.0f * m_fl.x / W;
        r(0, 2) = W - 2.0f * m_pp.x;
        r(1, 1) = 2.0f * m_fl.y / H;
        r(1, 2) = H - 2.0f * m_pp.y;
        r(2, 2) = -(zf + zn) / (zf - zn);
        r(2, 3) = -2.0f * zf * zn / (zf - zn);
        r(3, 2) = -1.0f;
        return r;
    }

    // DepthSensorParams.cpp:65-80
    static DepthSensorParams KinectParams(KinectDepthSensorResolution resolution, KinectDepthSensorMode mode)
    {
        const bool lo = resolution == KinectDepthSensorResolution320x240;
        const util::int2 res = lo ? util::int2(320, 240) : util::int2(640, 480);
        const float fl = lo ? 285.63f : 571.26f;
        const util::float2 range = mode == KinectDepthSensorModeNear ? util::float2(0.4f, 3.0f)
                                                                     : util::float2(0.8f, 4.0f);
        return DepthSensorParams(res, util::float2(fl), util::float2(res.x * 0.5f, res.y * 0.5f), range);
    }

    // the C ABI's view of the same camera
    kifi_intrinsics Intrinsics() const
    {
        kifi_intrinsics c;
        c.width = m_res.x;
        c.height = m_res.y;
        c.fx = m_fl.x;
        c.fy = m_fl.y;
        c.cx = m_pp.x;
        c.cy = m_pp.y;
        c.znear = m_range.x;
        c.zfar = m_range.y;
        return c;
    }

private:
    util::int2 m_res;
    util::float2 m_fl, m_pp, m_range;
};

} // namespace kifi
