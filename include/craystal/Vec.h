// Vec.h - the math types of the reference's public API (Core/Vec.h:8-63, where
// they are glm 1.0.1 aliases), as plain structs with the same member names:
// .x/.y/.z (.r/.g/.b, .s/.t/.p) and operator[], component-wise arithmetic and
// the glm functions the host code uses (dot, cross, normalize, length).
// Default construction zero-initialises.  The structs are standard layout
// and tightly packed (Float3 = 12 bytes), so &v[0].x of a std::vector<Float3>
// is a flat float array, which is how they cross the C-ABI.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>

namespace CRay {

using Float = float;
using UInt = uint32_t;

template <class T>
struct TVec2 {
    union { T x, r, s; };
    union { T y, g, t; };
    constexpr TVec2() : x(0), y(0) {}
    constexpr explicit TVec2(T v) : x(v), y(v) {}
    constexpr TVec2(T a, T b) : x(a), y(b) {}
    T& operator[](size_t i) { return i == 0 ? x : y; }
    const T& operator[](size_t i) const { return i == 0 ? x : y; }
    static constexpr int length() { return 2; }
};

template <class T>
struct TVec3 {
    union { T x, r, s; };
    union { T y, g, t; };
    union { T z, b, p; };
    constexpr TVec3() : x(0), y(0), z(0) {}
    constexpr explicit TVec3(T v) : x(v), y(v), z(v) {}
    constexpr TVec3(T a, T b, T c) : x(a), y(b), z(c) {}
    T& operator[](size_t i) { return i == 0 ? x : (i == 1 ? y : z); }
    const T& operator[](size_t i) const { return i == 0 ? x : (i == 1 ? y : z); }
    static constexpr int length() { return 3; }
};

using Float2 = TVec2<float>;
using Float3 = TVec3<float>;
using UInt2 = TVec2<uint32_t>;
using UInt3 = TVec3<uint32_t>;
using Int2 = TVec2<int32_t>;
using Int3 = TVec3<int32_t>;

static_assert(sizeof(Float2) == 8 && sizeof(Float3) == 12 && sizeof(UInt2) == 8,
              "vector types must be tightly packed");

// Component-wise arithmetic (glm's operators).
#define CRAY_VEC_OPS(V, N)                                                            \
    template <class T>                                                                \
    constexpr V<T> operator+(const V<T>& a, const V<T>& b) { return apply(a, b, [](T u, T v) { return u + v; }); } \
    template <class T>                                                                \
    constexpr V<T> operator-(const V<T>& a, const V<T>& b) { return apply(a, b, [](T u, T v) { return u - v; }); } \
    template <class T>                                                                \
    constexpr V<T> operator*(const V<T>& a, const V<T>& b) { return apply(a, b, [](T u, T v) { return u * v; }); } \
    template <class T>                                                                \
    constexpr V<T> operator/(const V<T>& a, const V<T>& b) { return apply(a, b, [](T u, T v) { return u / v; }); } \
    template <class T>                                                                \
    constexpr V<T> operator*(const V<T>& a, T s) { return a * V<T>(s); }             \
    template <class T>                                                                \
    constexpr V<T> operator*(T s, const V<T>& a) { return V<T>(s) * a; }             \
    template <class T>                                                                \
    constexpr V<T> operator/(const V<T>& a, T s) { return a / V<T>(s); }             \
    template <class T>                                                                \
    constexpr V<T> operator+(const V<T>& a, T s) { return a + V<T>(s); }             \
    template <class T>                                                                \
    constexpr V<T> operator-(const V<T>& a, T s) { return a - V<T>(s); }             \
    template <class T>                                                                \
    constexpr V<T> operator-(const V<T>& a) { return V<T>(T(0)) - a; }               \
    template <class T, class U>                                                       \
    V<T>& operator+=(V<T>& a, const U& b) { return a = a + b; }                       \
    template <class T, class U>                                                       \
    V<T>& operator-=(V<T>& a, const U& b) { return a = a - b; }                       \
    template <class T, class U>                                                       \
    V<T>& operator*=(V<T>& a, const U& b) { return a = a * b; }                       \
    template <class T, class U>                                                       \
    V<T>& operator/=(V<T>& a, const U& b) { return a = a / b; }                       \
    template <class T>                                                                \
    constexpr bool operator==(const V<T>& a, const V<T>& b) {                         \
        for (int i = 0; i < N; ++i)                                                   \
            if (!(a[i] == b[i])) return false;                                        \
        return true;                                                                  \
    }                                                                                 \
    template <class T>                                                                \
    constexpr bool operator!=(const V<T>& a, const V<T>& b) { return !(a == b); }

template <class T, class F>
constexpr TVec2<T> apply(const TVec2<T>& a, const TVec2<T>& b, F f) {
    return TVec2<T>(f(a.x, b.x), f(a.y, b.y));
}
template <class T, class F>
constexpr TVec3<T> apply(const TVec3<T>& a, const TVec3<T>& b, F f) {
    return TVec3<T>(f(a.x, b.x), f(a.y, b.y), f(a.z, b.z));
}

CRAY_VEC_OPS(TVec2, 2)
CRAY_VEC_OPS(TVec3, 3)
#undef CRAY_VEC_OPS

// glm::dot / cross / length / normalize, in glm 1.0.1's order of operations.
inline float dot(const Float3& a, const Float3& b) {
    Float3 t = a * b;
    return (t.x + t.y) + t.z;
}
inline float dot(const Float2& a, const Float2& b) { return a.x * b.x + a.y * b.y; }
inline Float3 cross(const Float3& a, const Float3& b) {
    return Float3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
inline float length(const Float3& v) { return std::sqrt(dot(v, v)); }
inline Float3 normalize(const Float3& v) { return v * (1.0f / std::sqrt(dot(v, v))); }

}  // namespace CRay
