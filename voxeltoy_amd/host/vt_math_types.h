/*
 * vt_math_types.h - the few small vector, box and matrix types the host
 * classes need (the reference uses Imath for these: V2f, V2i, V3f, V3i,
 * Box3f, M44f).  Float arithmetic is written out in the order the
 * reference's Imath calls evaluate it.
 */
#pragma once

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <utility>

namespace voxeltoy {

/* Vector types of other libraries (Imath::V2i, Imath::V3f, ...) convert to
 * and from these by their x, y (and z) members, so host code written
 * against Imath-typed fields can assign its own vectors to them:
 * has_xy<T> / has_xyz<T> select the two- and three-component types. */
namespace detail {
template <typename T, typename = void>
struct has_z : std::false_type {};
template <typename T>
struct has_z<T, std::void_t<decltype(std::declval<const T&>().z)>> : std::true_type {};
template <typename T, typename = void>
struct has_xy : std::false_type {};
template <typename T>
struct has_xy<T, std::void_t<decltype(std::declval<const T&>().x), decltype(std::declval<const T&>().y)>>
    : std::true_type {};
}  // namespace detail
template <typename T>
using enable_if_vec2 = std::enable_if_t<detail::has_xy<T>::value && !detail::has_z<T>::value, int>;
template <typename T>
using enable_if_vec3 = std::enable_if_t<detail::has_xy<T>::value && detail::has_z<T>::value, int>;

struct V2f {
    float x = 0.0f, y = 0.0f;
    V2f() = default;
    V2f(float a, float b) : x(a), y(b) {}
    explicit V2f(float a) : x(a), y(a) {}
    template <typename T, enable_if_vec2<T> = 0>
    V2f(const T& v) : x((float)v.x), y((float)v.y) {}
};

struct V2i {
    int x = 0, y = 0;
    V2i() = default;
    V2i(int a, int b) : x(a), y(b) {}
    template <typename T, enable_if_vec2<T> = 0>
    V2i(const T& v) : x((int)v.x), y((int)v.y) {}
};

struct V3f {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    V3f() = default;
    V3f(float a, float b, float c) : x(a), y(b), z(c) {}
    explicit V3f(float a) : x(a), y(a), z(a) {}
    template <typename T, enable_if_vec3<T> = 0>
    V3f(const T& v) : x((float)v.x), y((float)v.y), z((float)v.z) {}
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    V3f operator+(const V3f& o) const { return V3f(x + o.x, y + o.y, z + o.z); }
    V3f operator-(const V3f& o) const { return V3f(x - o.x, y - o.y, z - o.z); }
    V3f operator-() const { return V3f(-x, -y, -z); }
    V3f operator*(float s) const { return V3f(x * s, y * s, z * s); }
    V3f operator/(float s) const { return V3f(x / s, y / s, z / s); }
    float dot(const V3f& o) const { return x * o.x + y * o.y + z * o.z; }
    V3f cross(const V3f& o) const { return V3f(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x); }
    float length() const { return std::sqrt(dot(*this)); }
    V3f normalized() const
    {
        const float l = length();
        return l == 0.0f ? V3f(0.0f) : V3f(x / l, y / l, z / l);
    }
};
inline V3f operator*(float s, const V3f& v) { return v * s; }

struct V3i {
    int x = 0, y = 0, z = 0;
    V3i() = default;
    V3i(int a, int b, int c) : x(a), y(b), z(c) {}
    explicit V3i(int a) : x(a), y(a), z(a) {}
    template <typename T, enable_if_vec3<T> = 0>
    V3i(const T& v) : x((int)v.x), y((int)v.y), z((int)v.z) {}
    bool operator==(const V3i& o) const { return x == o.x && y == o.y && z == o.z; }
};

struct Box3f {
    V3f min = V3f(1e30f), max = V3f(-1e30f);   /* empty box */
    Box3f() = default;
    Box3f(const V3f& a, const V3f& b) : min(a), max(b) {}
    bool isEmpty() const { return max.x < min.x || max.y < min.y || max.z < min.z; }
    void extendBy(const V3f& p)
    {
        min = V3f(std::min(min.x, p.x), std::min(min.y, p.y), std::min(min.z, p.z));
        max = V3f(std::max(max.x, p.x), std::max(max.y, p.y), std::max(max.z, p.z));
    }
    V3f size() const { return isEmpty() ? V3f(0.0f) : max - min; }
    V3f center() const { return (max + min) / 2.0f; }
    /* index of the longest axis (ties -> lowest index) */
    int majorAxis() const
    {
        const V3f s = size();
        int m = 0;
        if (s.y > s[m]) m = 1;
        if (s.z > s[m]) m = 2;
        return m;
    }
};

/* Row-major 4x4 matrix: x[row][col]. */
struct M44f {
    float x[4][4] = { { 1, 0, 0, 0 }, { 0, 1, 0, 0 }, { 0, 0, 1, 0 }, { 0, 0, 0, 1 } };
    const float* data() const { return &x[0][0]; }
    float* data() { return &x[0][0]; }
};

}  // namespace voxeltoy
