// Device-side vector math, sampler, ray offset, slab and Moller-Trumbore
// tests for gfx950.
//
// Operation order follows the reference's glm-1.0.1 types (Core/Vec.h,
// Math/CRayMath.h): dot = (x*x' + y*y') + z*z', cross in glm's component
// order, normalize = v * (1/sqrt(dot)).  The library is compiled with
// -ffp-contract=off and HIP's default correctly rounded f32 division and
// square root, so every function here that uses only + - * / sqrt and
// compares returns the same bits as the host code it restates.
// min/max are comparison selects (std::min/std::max semantics), never
// v_min/v_max, because IEEE minNum differs on NaN.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "layout.h"

namespace crh {

#define CRH_D __device__ __forceinline__

constexpr float kEps = 1e-6f;
constexpr float kPi = 3.14159265358979323846f;
constexpr float k2Pi = 2.f * kPi;
constexpr float kInvPi = 1.f / kPi;
constexpr float kInv2Pi = 0.5f / kPi;
constexpr float kFltEps = 1.1920928955078125e-07f;

struct F3 {
    float x, y, z;
};
struct F2 {
    float x, y;
};

CRH_D F3 mk3(float a, float b, float c) { return F3{a, b, c}; }
CRH_D F3 ld3(const float* p) { return F3{p[0], p[1], p[2]}; }
CRH_D F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
CRH_D F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
CRH_D F3 operator*(F3 a, F3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
CRH_D F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
CRH_D F3 operator*(float s, F3 a) { return {s * a.x, s * a.y, s * a.z}; }
CRH_D F3 operator/(F3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
CRH_D F3 operator-(F3 a) { return {-a.x, -a.y, -a.z}; }
CRH_D float comp(F3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

CRH_D float dot(F3 a, F3 b) {
    float x = a.x * b.x, y = a.y * b.y, z = a.z * b.z;
    return (x + y) + z;
}
CRH_D float absDot(F3 a, F3 b) { return fabsf(dot(a, b)); }
CRH_D F3 cross(F3 a, F3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x,
            a.x * b.y - b.x * a.y};
}
CRH_D float length(F3 v) { return sqrtf(dot(v, v)); }
CRH_D F3 normalize(F3 v) { return v * (1.0f / sqrtf(dot(v, v))); }
CRH_D F3 reflect(F3 i, F3 n) { return i - n * dot(n, i) * 2.0f; }

// std::min(a, b) / std::max(a, b) select semantics.
CRH_D float smin(float a, float b) { return (b < a) ? b : a; }
CRH_D float smax(float a, float b) { return (a < b) ? b : a; }

// ---------------------------------------------------------------------------
// Sampler: PCG-style state step + RXS-M-XS output (Core/Sampler.cu:4-31).
// ---------------------------------------------------------------------------
struct Sampler {
    uint32_t state, inc;
    CRH_D Sampler(uint32_t x, uint32_t y, uint32_t sampleIndex) {
        state = x * 2654435761u + y * 1597334677u + sampleIndex;
        inc = x * 1234567891u + y * 987654321u + 1u;
#pragma unroll
        for (int i = 0; i < 4; ++i) next();
    }
    CRH_D uint32_t next() {
        uint32_t old = state;
        state = old * 747796405u + inc;
        old = ((old >> ((old >> 28u) + 4u)) ^ old) * 277803737u;
        return (old >> 22u) ^ old;
    }
    CRH_D float next1D() { return float(next()) * (1.0f / 4294967296.0f); }
    CRH_D F2 next2D() {
        float a = next1D();
        float b = next1D();
        return {a, b};
    }
};

// ---------------------------------------------------------------------------
// Integer-ULP ray origin offset (Ray Tracing Gems ch.6; Math/Ray.cu:11-22).
// ---------------------------------------------------------------------------
CRH_D float offsetAxis(float p, float n) {
    int of = static_cast<int>(256.0f * n);
    float pi = __int_as_float(__float_as_int(p) + ((p < 0) ? -of : of));
    return fabsf(p) < (1.0f / 32.0f) ? p + (1.0f / 65535.0f) * n : pi;
}
CRH_D F3 offsetRayOrigin(F3 p, F3 n) {
    return {offsetAxis(p.x, n.x), offsetAxis(p.y, n.y), offsetAxis(p.z, n.z)};
}

// ---------------------------------------------------------------------------
// Rays
// ---------------------------------------------------------------------------
struct Ray {
    F3 o;
    float tMin;
    F3 d;
    float tMax;
};

// Per-ray constants for the slab test: invD = 1/d is what the reference
// recomputes per box (Math/AABB.cu:23).
struct RayPre {
    F3 o, d, inv;
    float tMin, tMax;
};
// Boxes a node's box test must take through boxHit: the node's flat-box bits,
// or both when the ray's tMin or tMax is NaN (see boxHitFast).
CRH_D uint32_t slowBoxes(uint32_t flat, const RayPre& r) {
    return flat | (__builtin_isunordered(r.tMin, r.tMax) ? 3u : 0u);
}
CRH_D RayPre prep(const Ray& r) {
    RayPre p;
    p.o = r.o;
    p.d = r.d;
    p.inv = {1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z};
    // Canonical (quiet) once here, so that the fmaxf / fminf of every box
    // test need not quiet them again; values and comparisons are unchanged.
    p.tMin = __builtin_canonicalizef(r.tMin);
    p.tMax = __builtin_canonicalizef(r.tMax);
    return p;
}

// One slab axis of AABB::intersect (Math/AABB.cu:10-33).  Returns false when
// the flat-box/parallel-ray case rejects.  Because tMin only grows and tMax
// only shrinks, evaluating all three axes before the single `tMin > tMax`
// test is equivalent to the reference's per-axis early out.
CRH_D bool slabAxis(float bmin, float bmax, float o, float d, float inv,
                    float& tMin, float& tMax) {
    if (fabsf(bmax - bmin) < kFltEps) {
        if (fabsf(d) < kFltEps) return !(fabsf(o - bmin) > kFltEps);
        float t = (bmin - o) / d;
        tMin = smax(tMin, t);
        tMax = smin(tMax, t);
        return true;
    }
    float t0 = (bmin - o) * inv;
    float t1 = (bmax - o) * inv;
    if (inv < 0.0f) {
        float s = t0;
        t0 = t1;
        t1 = s;
    }
    tMin = smax(tMin, t0);
    tMax = smin(tMax, t1);
    return true;
}

CRH_D bool boxHit(const float* bmin, const float* bmax, const RayPre& r,
                  float& hitT) {
    float tMin = r.tMin, tMax = r.tMax;
    bool ok = slabAxis(bmin[0], bmax[0], r.o.x, r.d.x, r.inv.x, tMin, tMax);
    ok &= slabAxis(bmin[1], bmax[1], r.o.y, r.d.y, r.inv.y, tMin, tMax);
    ok &= slabAxis(bmin[2], bmax[2], r.o.z, r.d.z, r.inv.z, tMin, tMax);
    if (!ok || tMin > tMax) return false;
    if (tMin > r.tMax || tMax < r.tMin) return false;
    hitT = tMin;
    return true;
}

// boxHit for a box with no flat axis (no |max - min| < FLT_EPSILON; the
// builder marks the boxes that have one) and a ray whose tMin and tMax are
// not NaN.  fmaxf / fminf (v_max_f32 / v_min_f32) stand for the reference's
// `t0 > tMin ? t0 : tMin` selects: while the accumulator is not NaN both
// give the other operand for a NaN t0 / t1 (0 * inf on a slab plane) and the
// larger / smaller one otherwise; they can differ only in the sign of a zero
// tie, which none of the comparisons the traversal makes can see.  The
// reference's final range test (tMin > ray tMax or tMax < ray tMin) cannot
// reject here: tMin starts at the ray's tMin and only grows, tMax starts at
// its tMax and only shrinks, none of them NaN, so tMin <= tMax implies
// ray tMin <= tMin <= tMax <= ray tMax.
CRH_D bool boxHitFast(const float* bmin, const float* bmax, const RayPre& r, float& hitT) {
    const float o[3] = {r.o.x, r.o.y, r.o.z};
    const float inv[3] = {r.inv.x, r.inv.y, r.inv.z};
    float tMin = r.tMin, tMax = r.tMax;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float t0 = (bmin[k] - o[k]) * inv[k];
        float t1 = (bmax[k] - o[k]) * inv[k];
        const bool sw = inv[k] < 0.0f;
        tMin = fmaxf(tMin, sw ? t1 : t0);
        tMax = fminf(tMax, sw ? t0 : t1);
    }
    if (tMin > tMax) return false;
    hitT = tMin;
    return true;
}

// Moller-Trumbore on a GTri (Scene/Triangle.cu:40-75).
CRH_D bool triHit(F3 v0, F3 e1, F3 e2, const RayPre& r, float& u, float& v,
                  float& t) {
    F3 p = cross(r.d, e2);
    float det = dot(e1, p);
    if (fabsf(det) < 1e-6f) return false;
    float inv = 1.0f / det;
    F3 tv = r.o - v0;
    u = dot(tv, p) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    F3 q = cross(tv, e1);
    v = dot(r.d, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = dot(e2, q) * inv;
    if (t <= 0.0f) return false;
    return true;
}

}  // namespace crh
