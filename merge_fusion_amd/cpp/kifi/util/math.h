// kifi/util/math.h -- the value types the kifi C++ API passes (reference:
// KinectFission/kifi/util/math.h:15-80 for the types, :362-377 for the products,
// :615-625 for invert_transform).  Same names, members and storage as the
// reference's, so caller code compiles unchanged; arithmetic is plain binary32
// in the reference's operation order, without SSE types.
#pragma once

#include <cmath>

namespace kifi {
namespace util {

struct int2 {
    int x, y;
    int2() : x(0), y(0) {}
    int2(int s) : x(s), y(s) {}
    int2(int x_, int y_) : x(x_), y(y_) {}
};

struct float2 {
    float x, y;
    float2() : x(0), y(0) {}
    float2(float s) : x(s), y(s) {}
    float2(float x_, float y_) : x(x_), y(y_) {}
};

struct float3 {
    float x, y, z;
    float3() : x(0), y(0), z(0) {}
    explicit float3(float s) : x(s), y(s), z(s) {}
    float3(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
};

struct float4 {
    float x, y, z, w;
    float4() : x(0), y(0), z(0), w(0) {}
    explicit float4(float s) : x(s), y(s), z(s), w(s) {}
    float4(float x_, float y_, float z_, float w_) : x(x_), y(y_), z(z_), w(w_) {}
    float4(float3 v, float w_) : x(v.x), y(v.y), z(v.z), w(w_) {}
    float3 xyz() const { return float3(x, y, z); }
};

inline float4 operator+(float4 u, float4 v) { return float4(u.x + v.x, u.y + v.y, u.z + v.z, u.w + v.w); }
inline float4 operator-(float4 u, float4 v) { return float4(u.x - v.x, u.y - v.y, u.z - v.z, u.w - v.w); }
inline float4 operator*(float4 u, float4 v) { return float4(u.x * v.x, u.y * v.y, u.z * v.z, u.w * v.w); }
inline float4 operator*(float4 u, float s) { return float4(u.x * s, u.y * s, u.z * s, u.w * s); }

// Column storage: element (row r, column c) is component r of col<c>; a
// float4x4 is 16 contiguous floats, m[r + 4c].
struct float4x4 {
    float4 col0, col1, col2, col3;

    float4x4() {}
    explicit float4x4(float s) : col0(s), col1(s), col2(s), col3(s) {}
    float4x4(float m00, float m01, float m02, float m03,
             float m10, float m11, float m12, float m13,
             float m20, float m21, float m22, float m23,
             float m30, float m31, float m32, float m33)
        : col0(m00, m10, m20, m30), col1(m01, m11, m21, m31),
          col2(m02, m12, m22, m32), col3(m03, m13, m23, m33) {}
    // src in row-major order
    explicit float4x4(const float *src)
        : float4x4(src[0], src[1], src[2], src[3], src[4], src[5], src[6], src[7],
                   src[8], src[9], src[10], src[11], src[12], src[13], src[14], src[15]) {}

    float *data() { return &col0.x; }
    const float *data() const { return &col0.x; }
    float &operator()(int r, int c) { return data()[r + 4 * c]; }
    const float &operator()(int r, int c) const { return data()[r + 4 * c]; }

    static float4x4 identity()
    {
        return float4x4(1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1);
    }
};
static_assert(sizeof(float4x4) == 64, "float4x4 must be 16 packed floats");
static_assert(sizeof(float3) == 12, "float3 must be 3 packed floats");

// (col0*x + col1*y) + (col2*z + col3*w)
inline float4 operator*(const float4x4 &m, float4 v)
{
    return (m.col0 * v.x + m.col1 * v.y) + (m.col2 * v.z + m.col3 * v.w);
}

inline float4x4 operator*(const float4x4 &m, const float4x4 &n)
{
    float4x4 r;
    r.col0 = m * n.col0;
    r.col1 = m * n.col1;
    r.col2 = m * n.col2;
    r.col3 = m * n.col3;
    return r;
}

inline float4 homogenize(float4 v) { return float4(v.x / v.w, v.y / v.w, v.z / v.w, 1.0f); }

// util::transpose (math.h:627-630)
inline void transpose(float4x4 &m)
{
    float *a = m.data();
    for (int r = 0; r < 4; ++r)
        for (int c = r + 1; c < 4; ++c) {
            const float t = a[r + 4 * c];
            a[r + 4 * c] = a[c + 4 * r];
            a[c + 4 * r] = t;
        }
}

// In-place inverse of a rigid transform: [R | t] -> [R^T | -R^T t]
inline void invert_transform(float4x4 &m)
{
    const float4 t(-m.col3.x, -m.col3.y, -m.col3.z, 1.0f);
    m.col3 = float4(0.0f, 0.0f, 0.0f, 1.0f);
    float s;
    s = m(0, 1); m(0, 1) = m(1, 0); m(1, 0) = s;
    s = m(0, 2); m(0, 2) = m(2, 0); m(2, 0) = s;
    s = m(1, 2); m(1, 2) = m(2, 1); m(2, 1) = s;
    m.col3 = m * t;
}

} // namespace util
} // namespace kifi
