// kifi/util/vector2d.h -- row-major 2D array (reference: KinectFission/kifi/util/vector2d.h:11-99):
// a std::vector with a width; element (x, y) at y * width + x.
#pragma once

#include <cstddef>
#include <vector>

namespace kifi {
namespace util {

template <typename T, class Alloc = std::allocator<T>>
class vector2d : public std::vector<T, Alloc> {
public:
    vector2d(const Alloc &a = Alloc()) : std::vector<T, Alloc>(a), m_width(0) {}
    vector2d(std::size_t width, std::size_t height, const T &val = T(), const Alloc &a = Alloc())
        : std::vector<T, Alloc>(width * height, val, a), m_width(width) {}

    std::size_t width() const { return m_width; }
    std::size_t height() const { return m_width == 0 ? 0 : this->size() / m_width; }

    T &operator()(std::size_t x, std::size_t y) { return (*this)[y * m_width + x]; }
    const T &operator()(std::size_t x, std::size_t y) const { return (*this)[y * m_width + x]; }

    void resize(std::size_t width, std::size_t height, const T &val = T())
    {
        std::vector<T, Alloc>::resize(width * height, val);
        m_width = width;
    }

private:
    std::size_t m_width;
};

} // namespace util
} // namespace kifi
