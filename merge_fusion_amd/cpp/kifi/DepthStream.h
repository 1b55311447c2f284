// kifi/DepthStream.h -- reader of .depth files (reference: KinectFission/kifi/
// DepthStream.h:11-45, DepthStream.cpp:15-102; format: poly2depth/README.txt).
//
// Layout, little endian: a 15-byte magic, int version; version 2 adds int
// width, height and texel type (0 = uint16 mm, 1 = binary32 metres); version 1
// is 640 x 480 uint16.  Then int frame count, and per frame 16 floats of view
// matrix followed by width x height texels, row-major.
//
// The reference reads the 16 floats through float4x4's row-major constructor
// and transposes once more (the file stores row-vector matrices), so the view
// is the 16 floats taken as column storage, which is what this reader returns.
//
// NextFrame(vector2d<float>&) gives metres as the reference does (uint16 texels
// as mm * 0.001f, DepthStream.cpp:84-91); NextFrame(vector2d<uint16_t>&) gives the
// raw millimetres of a uint16 file for kifi_integrate, with no conversion.
// Where the reference would read past the end of a file whose header promises
// more frames than it holds (it does not check the stream), NextFrame returns
// false.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "kifi/util/math.h"
#include "kifi/util/vector2d.h"

namespace kifi {

class DepthStream {
public:
    explicit DepthStream(const std::string &fileName) : m_file(fileName.c_str(), std::ifstream::binary)
    {
        char magic[15];
        int version = 0;
        if (!m_file.read(magic, 15) || !read_int(version))
            return;
        if (version == 1) {
            m_frameWidth = 640;
            m_frameHeight = 480;
            m_texelType = SHORT;
        } else if (version == 2) {
            int t = 0;
            if (!read_int(m_frameWidth) || !read_int(m_frameHeight) || !read_int(t) || (t != SHORT && t != FLOAT))
                return;
            m_texelType = (TexelType)t;
        } else {
            return;
        }
        if (!read_int(m_nFrames) || m_frameWidth <= 0 || m_frameHeight <= 0 || m_nFrames < 0)
            return;
        m_ok = true;
    }
    DepthStream(const DepthStream &) = delete;
    DepthStream &operator=(const DepthStream &) = delete;

    bool NextFrame(util::vector2d<float> &outFrame, util::float4x4 &outView)
    {
        if (!next_view(outView))
            return false;
        outFrame.resize(m_frameWidth, m_frameHeight);
        if (m_texelType == FLOAT) {
            if (!read_bytes(outFrame.data(), outFrame.size() * 4))
                return false;
        } else {
            m_raw.resize((std::size_t)m_frameWidth * m_frameHeight);
            if (!read_bytes(m_raw.data(), m_raw.size() * 2))
                return false;
            for (std::size_t i = 0; i < m_raw.size(); ++i)
                outFrame[i] = (float)m_raw[i] * 0.001f;   // DepthStream.cpp:89-90
        }
        ++m_iFrame;
        return true;
    }

    // millimetre texels of a uint16 file as they are (false for a float file)
    bool NextFrame(util::vector2d<uint16_t> &outFrameMM, util::float4x4 &outView)
    {
        if (m_texelType != SHORT || !next_view(outView))
            return false;
        outFrameMM.resize(m_frameWidth, m_frameHeight);
        if (!read_bytes(outFrameMM.data(), outFrameMM.size() * 2))
            return false;
        ++m_iFrame;
        return true;
    }

    int FrameCount() const { return m_nFrames; }
    int FrameWidth() const { return m_frameWidth; }
    int FrameHeight() const { return m_frameHeight; }
    bool IsFloat() const { return m_texelType == FLOAT; }
    bool Good() const { return m_ok; }

private:
    enum TexelType { SHORT = 0, FLOAT = 1 };

    bool read_int(int &v)
    {
        int32_t t;
        if (!m_file.read(reinterpret_cast<char *>(&t), 4))
            return false;
        v = t;
        return true;
    }
    bool read_bytes(void *dst, std::size_t n) { return (bool)m_file.read(reinterpret_cast<char *>(dst), (std::streamsize)n); }
    bool next_view(util::float4x4 &outView)
    {
        if (!m_ok || m_iFrame >= m_nFrames)
            return false;
        float view[16];
        if (!read_bytes(view, 64))
            return false;
        outView = util::float4x4(view);   // row-major constructor ...
        util::transpose(outView);         // ... and the row-vector convention: column storage = the file's floats
        return true;
    }

    std::ifstream m_file;
    TexelType m_texelType = SHORT;
    int m_frameWidth = 0, m_frameHeight = 0;
    int m_nFrames = 0, m_iFrame = 0;
    bool m_ok = false;
    std::vector<uint16_t> m_raw;
};

} // namespace kifi
