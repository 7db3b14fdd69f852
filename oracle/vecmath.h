// ORACLE - TEST INFRASTRUCTURE ONLY.  Nothing in the product links this.
//
// glm-1.0.1-semantics vector shim for the CPU restatement of the reference.
// The reference's math types are glm aliases (Core/Vec.h:8-63); the
// formulas below follow glm's published source so that operation order
// (and therefore every rounding) matches:
//   dot(a,b)     = (a.x*b.x + a.y*b.y) + a.z*b.z         (tmp = a*b; x+y+z)
//   cross(x,y)   = (x.y*y.z - y.y*x.z, x.z*y.x - y.z*x.x, x.x*y.y - y.x*x.y)
//   normalize(v) = v * (1 / sqrt(dot(v,v)))
//   length(v)    = sqrt(dot(v,v))
//   reflect(I,N) = I - N * dot(N,I) * 2
//   min/max      = comparison selects ((y < x) ? y : x, (x < y) ? y : x)
//   v / s        = (v.x / s, v.y / s, v.z / s)   (true division)
// Compiled with -ffp-contract=off so no FMA is formed.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

struct V2 {
    float x = 0.f, y = 0.f;
    V2() = default;
    V2(float a, float b) : x(a), y(b) {}
    float operator[](int i) const { return i == 0 ? x : y; }
};

struct V3 {
    float x = 0.f, y = 0.f, z = 0.f;
    V3() = default;
    explicit V3(float s) : x(s), y(s), z(s) {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};

inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }

inline float dot(V3 a, V3 b) {
    V3 t = a * b;
    return t.x + t.y + t.z;
}
inline float absDot(V3 a, V3 b) { return std::fabs(dot(a, b)); }
inline V3 cross(V3 x, V3 y) {
    return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x,
            x.x * y.y - y.x * x.y};
}
inline float length(V3 v) { return std::sqrt(dot(v, v)); }
inline V3 normalize(V3 v) { return v * (1.0f / std::sqrt(dot(v, v))); }
inline V3 reflect(V3 I, V3 N) { return I - N * dot(N, I) * 2.0f; }

// std::min / std::max / glm::min / glm::max comparison semantics.
inline float fmin_std(float a, float b) { return (b < a) ? b : a; }
inline float fmax_std(float a, float b) { return (a < b) ? b : a; }
inline V3 vmin(V3 a, V3 b) {
    return {fmin_std(a.x, b.x), fmin_std(a.y, b.y), fmin_std(a.z, b.z)};
}
inline V3 vmax(V3 a, V3 b) {
    return {fmax_std(a.x, b.x), fmax_std(a.y, b.y), fmax_std(a.z, b.z)};
}

inline int floatAsInt(float v) {
    int i;
    std::memcpy(&i, &v, 4);
    return i;
}
inline float intAsFloat(int v) {
    float f;
    std::memcpy(&f, &v, 4);
    return f;
}

// x86-64 cvttss2si: NaN and out-of-range inputs give INT_MIN.
inline int cvt_trunc_x86(float x) {
    if (!(x >= -2147483648.0f && x < 2147483648.0f)) return INT32_MIN;
    return static_cast<int>(x);
}

constexpr float kEps = 1e-6f;                     // Math/MathDefs.h:12
constexpr float kPi = 3.14159265358979323846f;    // :13
constexpr float k2Pi = 2.f * kPi;                 // :14
constexpr float kInvPi = 1.f / kPi;               // :15
constexpr float kInv2Pi = 0.5f / kPi;             // :16
constexpr float kFltInf = __builtin_inff();       // :19
constexpr float kFltEps = 1.1920928955078125e-07f;  // :20 (FLT_EPSILON)

// RGB Spectrum (Core/Spectrum.h:15-233, 237-276), three channels.
struct Spec {
    float v[3] = {0.f, 0.f, 0.f};
    Spec() = default;
    explicit Spec(float s) { v[0] = v[1] = v[2] = s; }
    Spec(float a, float b, float c) { v[0] = a; v[1] = b; v[2] = c; }
    float maxValue() const {
        float m = v[0];
        for (int i = 1; i < 3; ++i) m = fmax_std(m, v[i]);
        return m;
    }
    float averageValue() const {
        float s = 0.0f;
        for (int i = 0; i < 3; ++i) s += v[i];
        return s / 3.0f;
    }
    bool hasNaN() const {
        return std::isnan(v[0]) || std::isnan(v[1]) || std::isnan(v[2]);
    }
    bool hasInf() const {
        return std::isinf(v[0]) || std::isinf(v[1]) || std::isinf(v[2]);
    }
};
inline Spec operator+(const Spec& a, const Spec& b) {
    return {a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]};
}
inline Spec operator-(const Spec& a, const Spec& b) {
    return {a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]};
}
inline Spec operator*(const Spec& a, const Spec& b) {
    return {a.v[0] * b.v[0], a.v[1] * b.v[1], a.v[2] * b.v[2]};
}
inline Spec operator*(const Spec& a, float s) {
    return {a.v[0] * s, a.v[1] * s, a.v[2] * s};
}
inline Spec operator*(float s, const Spec& a) {
    return {s * a.v[0], s * a.v[1], s * a.v[2]};
}
inline Spec operator/(const Spec& a, float s) {
    return {a.v[0] / s, a.v[1] / s, a.v[2] / s};
}

}  // namespace orc
