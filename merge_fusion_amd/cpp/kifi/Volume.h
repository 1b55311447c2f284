// kifi/Volume.h -- the sparse TSDF volume (reference: KinectFission/kifi/Volume.h:15-48,
// Volume.cpp:12-93).  The voxels live in HBM as two sorted arrays (keys, values);
// Data() copies them out as the reference's flat_map view (keys_cbegin /
// values_cbegin, util/flat_map.h:86-121).  Beyond the reference: 64-bit keys for
// resolutions above 1024 and the binary16 + 8-bit-weight voxel layout.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "kifi/Voxel.h"
#include "kifi/detail.h"
#include "kifi/util/math.h"

namespace kifi {

// Host snapshot of a volume's (key, voxel) pairs in key order.
class VolumeData {
public:
    std::size_t size() const { return m_values.size(); }
    const unsigned *keys_cbegin() const { return m_keys.data(); }
    const unsigned *keys_cend() const { return m_keys.data() + m_keys.size(); }
    const Voxel *values_cbegin() const { return m_values.data(); }
    const Voxel *values_cend() const { return m_values.data() + m_values.size(); }
    // all key widths: z << 2b | y << b | x with b = 10 (32-bit) or 21 (64-bit)
    const std::vector<uint64_t> &keys64() const { return m_keys64; }

private:
    friend class Volume;
    std::vector<unsigned> m_keys;       // 32-bit volumes only
    std::vector<uint64_t> m_keys64;
    std::vector<Voxel> m_values;
};

class Volume {
public:
    Volume(int resolution, float sideLength, float truncationMargin, int keyBits = 32,
           int storage = KIFI_STORAGE_F32)
        : m_res(resolution), m_side(sideLength), m_trunc(truncationMargin)
    {
        kifi_volume_desc d;
        d.res = resolution;
        d.side = sideLength;
        d.trunc = truncationMargin;
        d.key_bits = keyBits;
        d.storage = storage;
        m_desc = d;
        detail::check(kifi_volume_create(detail::context(), &d, &m_h), "kifi_volume_create");
        m_owned = true;
    }
    ~Volume()
    {
        if (m_owned)
            kifi_volume_destroy(m_h);
    }
    Volume(const Volume &) = delete;
    Volume &operator=(const Volume &) = delete;

    int Resolution() const { return m_res; }
    float SideLength() const { return m_side; }
    float VoxelLength() const { return m_side / m_res; }
    float TruncationMargin() const { return m_trunc; }
    util::float4 Minimum() const { return util::float4(-0.5f * m_side, -0.5f * m_side, -0.5f * m_side, 1.0f); }
    util::float4 Maximum() const { return util::float4(0.5f * m_side, 0.5f * m_side, 0.5f * m_side, 1.0f); }
    // VoxelCenter(x, y, z) = index * vl + (vl / 2 + min)  (Volume.cpp:28-36, 68-75)
    util::float3 VoxelCenter(int x, int y, int z) const
    {
        const float vl = VoxelLength(), o = vl * 0.5f + Minimum().x;
        return util::float3(x * vl + o, y * vl + o, z * vl + o);
    }

    std::size_t Size() const
    {
        std::size_t n = 0;
        detail::check(kifi_volume_size(m_h, &n), "kifi_volume_size");
        return n;
    }
    void Clear() { detail::check(kifi_volume_clear(m_h), "kifi_volume_clear"); }

    VolumeData Data() const
    {
        VolumeData out;
        const std::size_t n = Size();
        std::size_t got = 0;
        out.m_keys64.resize(n);
        out.m_values.resize(n);
        if (m_desc.key_bits == 32) {
            out.m_keys.resize(n);
            if (m_desc.storage == KIFI_STORAGE_F32) {
                detail::check(kifi_volume_download(m_h, out.m_keys.data(), out.m_values.data(), n, &got),
                              "kifi_volume_download");
            } else {
                std::vector<uint32_t> raw(n);
                detail::check(kifi_volume_download(m_h, out.m_keys.data(), raw.data(), n, &got),
                              "kifi_volume_download");
                unpack_f16u8(raw, out.m_values);
            }
            for (std::size_t i = 0; i < n; ++i)
                out.m_keys64[i] = out.m_keys[i];
        } else if (m_desc.storage == KIFI_STORAGE_F32) {
            detail::check(kifi_volume_download(m_h, out.m_keys64.data(), out.m_values.data(), n, &got),
                          "kifi_volume_download");
        } else {
            std::vector<uint32_t> raw(n);
            detail::check(kifi_volume_download(m_h, out.m_keys64.data(), raw.data(), n, &got),
                          "kifi_volume_download");
            unpack_f16u8(raw, out.m_values);
        }
        return out;
    }

    const kifi_volume_desc &Desc() const { return m_desc; }
    kifi_volume *Handle() const { return m_h; }

    // checkpoint / resume (kifi_volume_save / kifi_volume_load; no reference counterpart)
    void Save(const char *path) const { detail::check(kifi_volume_save(m_h, path), "kifi_volume_save"); }
    static Volume Load(const char *path)
    {
        kifi_volume *h = nullptr;
        detail::check(kifi_volume_load(detail::context(), path, &h), "kifi_volume_load");
        // the description from the file's header: magic, version, res, side, trunc, key bits, storage
        kifi_volume_desc d{};
        char head[32];
        FILE *f = std::fopen(path, "rb");
        const bool ok = f && std::fread(head, 1, 32, f) == 32;
        if (f)
            std::fclose(f);
        if (!ok) {
            kifi_volume_destroy(h);
            std::fprintf(stderr, "kifi: cannot read %s\n", path);
            std::abort();
        }
        int32_t i;
        std::memcpy(&i, head + 12, 4); d.res = i;
        std::memcpy(&d.side, head + 16, 4);
        std::memcpy(&d.trunc, head + 20, 4);
        std::memcpy(&i, head + 24, 4); d.key_bits = i;
        std::memcpy(&i, head + 28, 4); d.storage = i;
        Volume v(h, d);
        v.m_owned = true;
        return v;
    }

    // A volume owned by the library (a pipeline's); not destroyed by this object.
    static Volume Borrow(kifi_volume *h, const kifi_volume_desc &d) { return Volume(h, d); }
    Volume(Volume &&o) noexcept
        : m_res(o.m_res), m_side(o.m_side), m_trunc(o.m_trunc), m_desc(o.m_desc), m_h(o.m_h), m_owned(o.m_owned)
    {
        o.m_owned = false;
    }

private:
    Volume(kifi_volume *h, const kifi_volume_desc &d)
        : m_res(d.res), m_side(d.side), m_trunc(d.trunc), m_desc(d), m_h(h), m_owned(false) {}

    static float f16_to_f32(uint16_t h)
    {
        const uint32_t s = (h >> 15) & 1u, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
        float v;
        if (e == 0)
            v = std::ldexp((float)m, -24);
        else if (e == 31)
            v = m ? NAN : INFINITY;
        else
            v = std::ldexp((float)(m | 0x400u), (int)e - 25);
        return s ? -v : v;
    }
    // binary16 distance | weight << 16  ->  Voxel(distance * weight, weight)
    static void unpack_f16u8(const std::vector<uint32_t> &raw, std::vector<Voxel> &out)
    {
        for (std::size_t i = 0; i < raw.size(); ++i) {
            const float w = (float)((raw[i] >> 16) & 0xffu);
            out[i] = Voxel(f16_to_f32((uint16_t)(raw[i] & 0xffffu)) * w, w);
        }
    }

    int m_res;
    float m_side, m_trunc;
    kifi_volume_desc m_desc{};
    kifi_volume *m_h = nullptr;
    bool m_owned = false;
};

} // namespace kifi
