// ORACLE - TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference's rendering hot path (Hyiker/CRaystal),
// used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
// the checker.  The product (libcrayhip.so) never links, loads or calls it.
//
// Every function cites the reference file:line it restates (paths relative to
// the reference's Source/CRaystal/).  The reference's integrator, traversal,
// shading and film are CUDA device code; there is no CPU render path in the
// reference, so this file *is* "the reference's CPU path" for timing.
//
// Build: g++ -O2 -ffp-contract=off -fno-fast-math -fopenmp (oracle/Makefile).
//
// Parity pins (tests/test_oracle_*.py):
//   Sampler            closed-form integer arithmetic + the reference's
//                      property tests (Tests/CoreTest.cpp:130-281)
//   Camera             reference KATs (Tests/CameraTest.cpp:16-50)
//   offsetRayOrigin    hand KATs of the RTG ch.6 formula
//   AABB / MT          textbook slab / Moller-Trumbore special cases, brute force
//   BVH build          invariants + brute-force argmin of the SAH split
//   Traversal          brute-force closest hit over all triangles
//   BSDFs              closed forms (Lambert, Schlick, GGX normalisation)
//   Integrator         furnace-style closed forms (emitter-only, white box)
//   Film               tent-filter closed form, constant-image KAT
//   Spheres            float64 roots, inside / range / miss cases, the (u, v)
//                      closed form, brute force over spheres + triangles
//   Textures           bilinear closed form, channel fill, saturating casts
//   RGB uplift         colour round trip through the CIE table, bounds,
//                      homogeneity, continuity across channel orderings
//   Principled eval    hand-derived closed forms (r = 1, normal incidence,
//                      grazing views) that fail for each of five planted
//                      mistakes (ORC_MUTANT)
//   CIE table          the standard observer's peaks and integrals; Y of a
//                      flat spectrum = (470/95) / (469/94)
//   Guided estimator   the sampler's actual density vs histograms; mixture
//                      vs the float64 model; a guide equal to the BSDF gives
//                      the plain weight; closed-form and plain-image means
// Parity unpinned at bit level: glm's exact operation order beyond the
// camera KAT tolerance and CUDA libdevice transcendental results.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include "../craystal_amd/csrc/cie_table.h"  // CIE 1931 table data (kXYZTable)
#include "../craystal_amd/csrc/rgb_basis.h"  // basis data of the RGB uplift
#include "../include/crayhip.h"
#include "vecmath.h"

#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

namespace orc {

// ---------------------------------------------------------------------------
// Sampler (Core/Sampler.cu:4-31)
// ---------------------------------------------------------------------------
struct Sampler {
    uint32_t state, inc;
    Sampler(uint32_t px, uint32_t py, uint32_t sampleIndex) {
        state = px * 2654435761u + py * 1597334677u + sampleIndex;  // :6
        inc = px * 1234567891u + py * 987654321u + 1u;              // :7
        for (int i = 0; i < 4; ++i) next();                         // :8-10
    }
    uint32_t next() {  // :21-27
        uint32_t old = state;
        state = old * 747796405u + inc;
        old = ((old >> ((old >> 28u) + 4u)) ^ old) * 277803737u;
        return (old >> 22u) ^ old;
    }
    float next1D() { return float(next()) * (1.0f / 4294967296.0f); }  // :29-31
    V2 next2D() {  // :17-19 (x drawn before y)
        float a = next1D();
        float b = next1D();
        return {a, b};
    }
};

// ---------------------------------------------------------------------------
// Ray / offset (Math/Ray.h:8-19, Math/Ray.cu:11-22)
// ---------------------------------------------------------------------------
struct Ray {
    V3 o;
    float tMin = 0.f;
    V3 d;
    float tMax = kFltInf;
};

V3 offsetRayOrigin(V3 p, V3 n) {
    const float kOrigin = 1.0f / 32.0f;
    const float kFltScale = 1.0f / 65535.0f;
    const float kIntScale = 256.0f;
    int ox = static_cast<int>(kIntScale * n.x);
    int oy = static_cast<int>(kIntScale * n.y);
    int oz = static_cast<int>(kIntScale * n.z);
    V3 pI(intAsFloat(floatAsInt(p.x) + ((p.x < 0) ? -ox : ox)),
          intAsFloat(floatAsInt(p.y) + ((p.y < 0) ? -oy : oy)),
          intAsFloat(floatAsInt(p.z) + ((p.z < 0) ? -oz : oz)));
    return V3(std::fabs(p.x) < kOrigin ? p.x + kFltScale * n.x : pI.x,
              std::fabs(p.y) < kOrigin ? p.y + kFltScale * n.y : pI.y,
              std::fabs(p.z) < kOrigin ? p.z + kFltScale * n.z : pI.z);
}

// ---------------------------------------------------------------------------
// AABB (Math/AABB.h:12-138, Math/AABB.cu:6-46)
// ---------------------------------------------------------------------------
struct AABB {
    V3 pMin{kFltInf, kFltInf, kFltInf};
    V3 pMax{-kFltInf, -kFltInf, -kFltInf};
    void include(V3 p) {
        pMin = vmin(pMin, p);
        pMax = vmax(pMax, p);
    }
    void include(const AABB& b) {
        pMin = vmin(pMin, b.pMin);
        pMax = vmax(pMax, b.pMax);
    }
    V3 diagonal() const { return pMax - pMin; }
};

bool aabbIntersect(const AABB& box, const Ray& ray, float& hitT) {
    float tMin = ray.tMin;
    float tMax = ray.tMax;
    for (int i = 0; i < 3; ++i) {
        if (std::fabs(box.pMax[i] - box.pMin[i]) < kFltEps) {  // :12
            if (std::fabs(ray.d[i]) < kFltEps) {                // :13
                if (std::fabs(ray.o[i] - box.pMin[i]) > kFltEps) return false;
                continue;
            }
            float t = (box.pMin[i] - ray.o[i]) / ray.d[i];  // :19
            tMin = fmax_std(tMin, t);
            tMax = fmin_std(tMax, t);
        } else {
            float invD = 1.0f / ray.d[i];  // :23
            float t0 = (box.pMin[i] - ray.o[i]) * invD;
            float t1 = (box.pMax[i] - ray.o[i]) * invD;
            if (invD < 0.0f) std::swap(t0, t1);
            tMin = fmax_std(tMin, t0);
            tMax = fmin_std(tMax, t1);
        }
        if (tMin > tMax) return false;  // :35
    }
    if (tMin > ray.tMax || tMax < ray.tMin) return false;  // :40
    hitT = tMin;
    return true;
}

// ---------------------------------------------------------------------------
// Scene storage (TriangleMeshManager, Scene/Triangle.cu:98-165;
// MaterialManager, Core/Material.cu:15-30)
// ---------------------------------------------------------------------------
struct NodeRef {  // BVHNode (Core/BVH.h:17-30), canonical export layout
    uint32_t tag, a, b;
    AABB bounds;
};

struct BVHOut {
    std::vector<NodeRef> nodes;
    std::vector<uint32_t> primOrder;  // bvhToMeshPrimitiveIndex
    AABB rootBounds;
    uint32_t root = 0;
    uint32_t maxDepth = 0, maxLeaf = 0;
    double buildMs = 0;
};

// CIE 1931 colour matching functions of spectral bin k: the reference's
// kXYZTable (Core/SpectrumTabData.h:7-55), with its step kXYZStep = 470/95
// (:57) and normalisation kCIENormalizationY (Core/Spectrum.h:286).  The data
// header is generated by craystal_amd/csrc/make_cie_table.py from the CIE's
// 1 nm table as the reference's Tools/spectrum.py resampled it; the values
// are pinned against the standard observer in tests/test_oracle_pins.py.
inline void cieXYZ(int k, double out[3]) {
    for (int j = 0; j < 3; ++j) out[j] = double(crh::kCieXYZ[j][k]);
}
// BroadSpectrum::toXYZ (Core/Spectrum.cpp:37-45): float accumulation in bin
// order, then * (kCIENormalizationY * kXYZStep).  Z reads z-bar (the
// reference reads x-bar, :40, a bug fixed here on purpose).
inline void spectrumToXYZ(const float* s, float out[3]) {
    float acc[3] = {0.0f, 0.0f, 0.0f};
    for (int i = 0; i < 95; ++i)
        for (int j = 0; j < 3; ++j) acc[j] += crh::kCieXYZ[j][i] * s[i];
    const float k = crh::kCieNormY * crh::kCieXYZStep;
    for (int j = 0; j < 3; ++j) out[j] = acc[j] * k;
}

struct Scene {
    std::vector<V3> pos, nrm;
    std::vector<V2> uv;
    std::vector<uint32_t> idx;       // global (offset) vertex indices
    std::vector<uint32_t> primMesh;  // primitive -> mesh
    std::vector<uint32_t> meshMat;   // mesh -> material id
    std::vector<crh_material> mats;
    std::vector<uint32_t> emissive;
    Spec env;
    BVHOut bvh;
    // Spectral transport (an addition; the reference transports RGB): per material 3 x 95 bins
    // (diffuse, specular, emission), their bin averages, the environment
    // spectrum and the film's per-lane CIE weights.
    bool spectral = false;
    std::vector<float> spectra, envSpec, avgKd, avgKs;
    float filmW[95][3] = {};
    // Spheres (SphereSOA, Scene/Sphere.h): centre and radius; sphereMat
    // empty = the reference's rule (the triangle with the same index).
    std::vector<V3> sphC;
    std::vector<float> sphR;
    std::vector<uint32_t> sphereMat;
    // Textures (Texture, Core/Texture.h): 4 floats per texel.
    struct Tex {
        uint32_t w, h;
        std::vector<float> px;
    };
    std::vector<Tex> textures;
};

// ---------------------------------------------------------------------------
// BVH build (Core/BVH.cu:88-338)
// ---------------------------------------------------------------------------
struct BuildTri {  // Core/BVH.h:73-78
    V3 centroid;
    AABB bounds;
    uint32_t triIndex;
};

struct Builder {
    std::vector<BuildTri> tris;
    std::vector<NodeRef> nodes;
    uint32_t numBins = 32, maxLeaf = 4, maxDepthCap = 20;
    uint32_t statMaxDepth = 0, statMaxLeaf = 0;

    // calcAABBSurfaceArea (:284-287): float products and sums, then * 2.0.
    static float surfaceArea(const AABB& b) {
        V3 d = b.diagonal();
        return float((d.x * d.y + d.x * d.z + d.y * d.z) * 2.0);
    }
    // evaluateSAH (:289-301)
    static float sah(const AABB& node, const AABB& l, const AABB& r, int nl,
                     int nr) {
        const float traversalCost = 1.0f, intersectionCost = 1.0f;
        float lsa = surfaceArea(l), rsa = surfaceArea(r),
              root = surfaceArea(node);
        return traversalCost +
               intersectionCost * (nl * lsa + nr * rsa) / root;
    }

    // findBestSplit (:214-282)
    void findBestSplit(const std::vector<uint32_t>& ind, uint32_t start,
                       uint32_t count, const AABB& nb, uint32_t& bestAxis,
                       float& bestSplit) {
        float bestCost = kFltInf;
        bestAxis = 0;
        bestSplit = 0.0f;
        const int NB = int(numBins);
        std::vector<AABB> binB(NB), lb(NB - 1), rb(NB - 1);
        std::vector<int> binC(NB), lc(NB - 1), rc(NB - 1);
        for (uint32_t axis = 0; axis < 3; ++axis) {
            std::fill(binB.begin(), binB.end(), AABB());
            std::fill(binC.begin(), binC.end(), 0);
            V3 ext = nb.diagonal();
            float scale = float(NB) / ext[axis];  // :226
            for (uint32_t i = start; i < start + count; ++i) {
                const BuildTri& t = tris[ind[i]];
                int b = std::min(
                    NB - 1,
                    cvt_trunc_x86((t.centroid[axis] - nb.pMin[axis]) * scale));
                b = std::clamp(b, 0, NB - 1);  // :235
                binB[b].include(t.bounds);
                binC[b]++;
            }
            AABB cur;
            int cc = 0;
            for (int i = 0; i < NB - 1; ++i) {  // :241-246
                cur.include(binB[i]);
                cc += binC[i];
                lb[i] = cur;
                lc[i] = cc;
            }
            AABB curR;
            int cr = 0;
            for (int i = NB - 1; i > 0; --i) {  // :250-255
                curR.include(binB[i]);
                cr += binC[i];
                rb[i - 1] = curR;
                rc[i - 1] = cr;
            }
            for (int i = 0; i < NB - 1; ++i) {  // :258-271
                float splitPos =
                    nb.pMin[axis] + float(i + 1) * ext[axis] / float(NB);
                float cost = sah(nb, lb[i], rb[i], lc[i], rc[i]);
                if (cost < bestCost) {
                    bestCost = cost;
                    bestAxis = axis;
                    bestSplit = splitPos;
                }
            }
        }
    }

    // buildRecursive (:150-212)
    uint32_t build(std::vector<uint32_t>& ind, uint32_t start, uint32_t count,
                   uint32_t depth, AABB& nb) {
        nb = AABB();
        for (uint32_t i = start; i < start + count; ++i)
            nb.include(tris[ind[i]].bounds);
        NodeRef node;
        node.bounds = nb;
        if (count <= maxLeaf || depth >= maxDepthCap) {
            node.tag = 0;
            node.a = start;
            node.b = count;
            nodes.push_back(node);
            statMaxLeaf = std::max(statMaxLeaf, count);
            return uint32_t(nodes.size() - 1);
        }
        uint32_t axis;
        float split;
        findBestSplit(ind, start, count, nb, axis, split);
        // libstdc++ std::partition (bidirectional two-pointer algorithm).
        auto mid = std::partition(
            ind.begin() + start, ind.begin() + start + count,
            [&](uint32_t id) { return tris[id].centroid[axis] < split; });
        uint32_t leftCount = uint32_t(std::distance(ind.begin() + start, mid));
        if (leftCount == 0 || leftCount == count) {  // :186-195
            node.tag = 0;
            node.a = start;
            node.b = count;
            nodes.push_back(node);
            statMaxLeaf = std::max(statMaxLeaf, count);
            return uint32_t(nodes.size() - 1);
        }
        AABB lb, rb;
        uint32_t l = build(ind, start, leftCount, depth + 1, lb);
        uint32_t r = build(ind, start + leftCount, count - leftCount,
                           depth + 1, rb);
        node.tag = 1;
        node.a = l;
        node.b = r;
        nodes.push_back(node);
        statMaxDepth = std::max(statMaxDepth, depth);
        return uint32_t(nodes.size() - 1);
    }
};

// BVH::build + gatherTriangles (:88-148) over a global triangle soup.
BVHOut buildBVH(const std::vector<V3>& pos, const std::vector<uint32_t>& idx,
                uint32_t bins, uint32_t maxLeaf, uint32_t maxDepth) {
    auto t0 = std::chrono::steady_clock::now();
    Builder b;
    b.numBins = bins ? bins : 32;
    b.maxLeaf = maxLeaf ? maxLeaf : 4;
    b.maxDepthCap = maxDepth ? maxDepth : 20;
    size_t ntri = idx.size() / 3;
    b.tris.resize(ntri);
    for (size_t i = 0; i < ntri; ++i) {
        V3 v0 = pos[idx[3 * i]], v1 = pos[idx[3 * i + 1]],
           v2 = pos[idx[3 * i + 2]];
        BuildTri& t = b.tris[i];
        t.centroid = (v0 + v1 + v2) / 3.0f;  // :138
        t.bounds = AABB();
        t.bounds.include(v0);
        t.bounds.include(v1);
        t.bounds.include(v2);
        t.triIndex = uint32_t(i);
    }
    std::vector<uint32_t> ind(ntri);
    std::iota(ind.begin(), ind.end(), 0u);
    b.nodes.reserve(2 * ntri);
    BVHOut out;
    out.root = b.build(ind, 0, uint32_t(ntri), 0, out.rootBounds);
    out.nodes = std::move(b.nodes);
    out.primOrder.resize(ntri);
    for (size_t i = 0; i < ntri; ++i) out.primOrder[i] = b.tris[ind[i]].triIndex;
    out.maxDepth = b.statMaxDepth;
    out.maxLeaf = b.statMaxLeaf;
    out.buildMs = std::chrono::duration<double, std::milli>(
                      std::chrono::steady_clock::now() - t0)
                      .count();
    return out;
}

// ---------------------------------------------------------------------------
// Triangle (Scene/Triangle.cu:9-96), intersectShape (Scene/Shape.h:17-37)
// ---------------------------------------------------------------------------
struct Vtx {
    V3 p, n;
    V2 uv;
};
struct Tri {
    Vtx v[3];
    V3 faceNormal() const {  // :9-16
        V3 e1 = v[1].p - v[0].p;
        V3 e2 = v[2].p - v[0].p;
        return normalize(cross(e1, e2));
    }
    Vtx interpolate(V3 w) const {  // :18-31
        Vtx r;
        r.p = v[0].p * w[0] + v[1].p * w[1] + v[2].p * w[2];
        r.n = v[0].n * w[0] + v[1].n * w[1] + v[2].n * w[2];
        r.uv = V2(v[0].uv.x * w[0] + v[1].uv.x * w[1] + v[2].uv.x * w[2],
                  v[0].uv.y * w[0] + v[1].uv.y * w[1] + v[2].uv.y * w[2]);
        return r;
    }
    float area() const {  // :33-37
        V3 a = v[0].p - v[1].p;
        V3 b = v[2].p - v[1].p;
        return length(cross(a, b)) / 2.f;
    }
};

Tri getTriangle(const Scene& s, uint32_t id) {  // :78-92
    Tri t;
    for (int i = 0; i < 3; ++i) {
        uint32_t vi = s.idx[3 * id + i];
        t.v[i].p = s.pos[vi];
        t.v[i].n = s.nrm[vi];
        t.v[i].uv = s.uv[vi];
    }
    return t;
}

// Moller-Trumbore (Scene/Triangle.cu:40-75).
bool triIntersect(const Scene& s, uint32_t id, const Ray& ray, float& u,
                  float& v, float& t) {
    V3 v0 = s.pos[s.idx[3 * id]], v1 = s.pos[s.idx[3 * id + 1]],
       v2 = s.pos[s.idx[3 * id + 2]];
    V3 e1 = v1 - v0;
    V3 e2 = v2 - v0;
    V3 p = cross(ray.d, e2);
    float det = dot(e1, p);
    if (std::fabs(det) < 1e-6f) return false;
    float inv = 1.0f / det;
    V3 tv = ray.o - v0;
    u = dot(tv, p) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    V3 q = cross(tv, e1);
    v = dot(ray.d, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = dot(e2, q) * inv;
    if (t <= 0.0) return false;
    return true;
}

struct RayHit {
    Ray ray;
    V2 bary;
    uint32_t prim = CRH_INVALID_PRIM;
    bool sphere = false;   // HitInfo::type == HitType::Sphere
    float hitT = kFltInf;  // Core/Hit.h:29
};

// intersectShape (Scene/Shape.h:17-37).
inline bool intersectShape(const Scene& s, uint32_t id, RayHit& rh) {
    float u = 0, v = 0, t = kFltInf;
    bool hit = triIntersect(s, id, rh.ray, u, v, t);
    if (!hit) return false;  // an unset hitT only ever leads to `false`
    if (t < rh.ray.tMin || t > rh.ray.tMax) return false;
    if (t < rh.hitT) {
        rh.bary = V2(u, v);
        rh.prim = id;
        rh.sphere = false;
        rh.hitT = t;
    }
    return true;
}

// SphereSOA::intersect (Scene/Sphere.cu:6-49): nearer root in range, else
// the farther; (u, v) from the unit normal at the hit point.
bool sphereIntersect(const Scene& s, uint32_t id, const Ray& ray, V2& uv, float& hitT) {
    V3 center = s.sphC[id];
    float radius = s.sphR[id];
    V3 oc = ray.o - center;
    float a = dot(ray.d, ray.d);
    float b = 2.0f * dot(oc, ray.d);
    float c = dot(oc, oc) - radius * radius;
    float disc = b * b - 4.0f * a * c;
    if (disc < 0) return false;
    float sq = std::sqrt(disc);
    float t0 = (-b - sq) / (2.0f * a);
    float t1 = (-b + sq) / (2.0f * a);
    if (t0 > t1) std::swap(t0, t1);
    if (t0 < ray.tMin || t0 > ray.tMax) {
        if (t1 < ray.tMin || t1 > ray.tMax) return false;
        t0 = t1;
    }
    hitT = t0;
    V3 hp = ray.o + ray.d * hitT;
    V3 n = normalize(hp - center);
    float theta = std::acos(-n.y);
    float phi = std::atan2(-n.z, n.x) + kPi;
    uv = V2(phi * (0.5f * kInvPi), theta * kInvPi);
    return true;
}
// intersectShape (Scene/Shape.h:17-37) on a sphere.
inline bool intersectSphereShape(const Scene& s, uint32_t id, RayHit& rh) {
    V2 uv;
    float t = kFltInf;
    if (!sphereIntersect(s, id, rh.ray, uv, t)) return false;
    if (t < rh.ray.tMin || t > rh.ray.tMax) return false;
    if (t < rh.hitT) {
        rh.bary = uv;
        rh.prim = id;
        rh.sphere = true;
        rh.hitT = t;
    }
    return true;
}

struct Counters {
    uint64_t nodes = 0, tris = 0;
};

// BVHData::intersect (Core/BVH.cu:13-84).  anyHit returns at the first
// reported hit; before that it visits exactly the nodes the closest-hit
// traversal visits, so the boolean equals intersectOcclusion's.
bool bvhIntersect(const Scene& s, RayHit& rh, Counters* cnt, bool anyHit) {
    const BVHOut& b = s.bvh;
    if (b.nodes.empty()) return false;
    float dummy;
    if (!aabbIntersect(b.rootBounds, rh.ray, dummy)) return false;
    struct Entry {
        uint32_t node;
        float tMin;
    };
    Entry stack[64];
    int sp = 0;
    stack[sp++] = {b.root, 0.0f};
    bool isHit = false;
    float closestT = rh.hitT;
    while (sp > 0) {
        Entry e = stack[--sp];
        const NodeRef& node = b.nodes[e.node];
        if (e.tMin > closestT) continue;
        if (node.tag == 0) {
            for (uint32_t i = 0; i < node.b; ++i) {
                uint32_t prim = b.primOrder[node.a + i];
                if (cnt) cnt->tris++;
                if (intersectShape(s, prim, rh)) {
                    isHit = true;
                    closestT = fmin_std(closestT, rh.hitT);
                    if (anyHit) return true;
                }
            }
        } else {
            if (cnt) cnt->nodes++;
            const NodeRef& L = b.nodes[node.a];
            const NodeRef& R = b.nodes[node.b];
            float lt = kFltInf, rt = kFltInf;
            bool hl = aabbIntersect(L.bounds, rh.ray, lt);
            bool hr = aabbIntersect(R.bounds, rh.ray, rt);
            if (hl && hr) {
                if (lt > rt) {
                    if (lt < closestT) stack[sp++] = {node.a, lt};
                    if (rt < closestT) stack[sp++] = {node.b, rt};
                } else {
                    if (rt < closestT) stack[sp++] = {node.b, rt};
                    if (lt < closestT) stack[sp++] = {node.a, lt};
                }
            } else if (hl && lt < closestT) {
                stack[sp++] = {node.a, lt};
            } else if (hr && rt < closestT) {
                stack[sp++] = {node.b, rt};
            }
        }
    }
    return isHit;
}

// SceneView::intersect / intersectOcclusion (Scene/Scene.cu:5-29): every
// sphere in index order, then the BVH over the triangles.  anyHit: the same
// boolean with an early exit.
bool sceneIntersect(const Scene& s, RayHit& rh, Counters* cnt, bool anyHit) {
    bool hit = false;
    for (uint32_t i = 0; i < uint32_t(s.sphR.size()); ++i)
        if (intersectSphereShape(s, i, rh)) {
            hit = true;
            if (anyHit) return true;
        }
    if (bvhIntersect(s, rh, cnt, anyHit)) hit = true;
    return hit;
}

// ---------------------------------------------------------------------------
// Frame / Intersection (Core/Frame.cu:6-31, Core/Intersection.cu:4-10,
// Scene/Scene.cu:31-58)
// ---------------------------------------------------------------------------
struct Frame {
    V3 T{1, 0, 0}, B{0, 1, 0}, N{0, 0, 1};
    static V3 tangentSafe(V3 n) {  // :13-20
        V3 t(0, 1, 0);
        if (1.0 - double(absDot(n, t)) < double(kEps)) t = V3(1, 0, 0);
        V3 bt = normalize(cross(n, t));
        return normalize(cross(bt, n));
    }
    explicit Frame(V3 n) {
        N = n;
        T = tangentSafe(n);
        B = cross(N, T);
    }
    Frame() = default;
    V3 toLocal(V3 w) const {  // mat3(T.x,B.x,N.x, T.y,...) * w
        return V3(T.x * w.x + T.y * w.y + T.z * w.z,
                  B.x * w.x + B.y * w.y + B.z * w.z,
                  N.x * w.x + N.y * w.y + N.z * w.z);
    }
    V3 toWorld(V3 l) const {  // mat3(T, B, N) * l
        return V3(T.x * l.x + B.x * l.y + N.x * l.z,
                  T.y * l.x + B.y * l.y + N.y * l.z,
                  T.z * l.x + B.z * l.y + N.z * l.z);
    }
};

struct Isect {
    V3 posW, faceN, viewW;
    V2 uv;
    bool front;
    Frame frame;
    V3 orientedFaceNormal() const { return front ? faceN : -faceN; }
};

Isect createIntersection(const Scene& s, const RayHit& rh) {
    Isect it;
    it.posW = rh.ray.o + rh.hitT * rh.ray.d;
    it.viewW = -normalize(rh.ray.d);
    V3 sn;
    if (rh.sphere) {  // HitType::Sphere (:45-49)
        it.faceN = normalize(it.posW - s.sphC[rh.prim]);
        sn = it.faceN;
        it.uv = rh.bary;
    } else {
        V3 w(1.f - rh.bary.x - rh.bary.y, rh.bary.x, rh.bary.y);  // Hit.h:20-23
        Tri tri = getTriangle(s, rh.prim);
        it.faceN = tri.faceNormal();
        Vtx vd = tri.interpolate(w);
        sn = normalize(vd.n);
        it.uv = vd.uv;
    }
    it.front = dot(it.viewW, it.faceN) > 0.0;
    it.frame = Frame(it.front ? sn : -sn);
    return it;
}

// TextureView::fetchTexel / sample (Core/Texture.cu:6-35).  The float ->
// uint32 casts follow the GPU's conversion (NaN and negatives to 0,
// saturating), which is what the reference's device code computes.
inline uint32_t gpuU32(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return uint32_t(f);
}
struct F4 {
    float v[4];
};
inline F4 lerpF4(const F4& a, const F4& b, float t) {  // Math/CRayMath.h:31-33
    F4 r;
    for (int i = 0; i < 4; ++i) r.v[i] = a.v[i] * (1.0f - t) + b.v[i] * t;
    return r;
}
F4 textureSample(const Scene::Tex& tx, V2 uv) {
    auto fetch = [&](uint32_t x, uint32_t y) {
        x = std::clamp<uint32_t>(x, 0, tx.w - 1);
        y = std::clamp<uint32_t>(y, 0, tx.h - 1);
        const float* p = tx.px.data() + 4 * (size_t(y) * tx.w + x);
        return F4{{p[0], p[1], p[2], p[3]}};
    };
    float px = uv.x * float(tx.w - 1);
    float py = uv.y * float(tx.h - 1);
    uint32_t x0 = gpuU32(std::floor(px)), y0 = gpuU32(std::floor(py));
    uint32_t x1 = std::min(x0 + 1, tx.w - 1), y1 = std::min(y0 + 1, tx.h - 1);
    float fx = px - float(x0), fy = py - float(y0);
    F4 c0 = lerpF4(fetch(x0, y0), fetch(x1, y0), fx);
    F4 c1 = lerpF4(fetch(x0, y1), fetch(x1, y1), fx);
    return lerpF4(c0, c1, fy);
}

// RGB -> 95-bin spectrum (crh_rgb_to_spectrum; Smits 1999): the flat white
// plus one secondary and one primary basis spectrum weighted by the sorted
// channel differences.  Returns the bins b[0..2] and the 95-bin average.
Spec upliftBins(float r, float g, float b, const uint32_t* bins, float& avg) {
    r = std::max(r, 0.0f);
    g = std::max(g, 0.0f);
    b = std::max(b, 0.0f);
    enum { C, M, Y, R, G, B };
    float white, w1, w2;
    int k1, k2;
    if (r <= g && r <= b) {
        white = r; k1 = C;
        if (g <= b) { w1 = g - r; k2 = B; w2 = b - g; } else { w1 = b - r; k2 = G; w2 = g - b; }
    } else if (g <= r && g <= b) {
        white = g; k1 = M;
        if (r <= b) { w1 = r - g; k2 = B; w2 = b - r; } else { w1 = b - g; k2 = R; w2 = r - b; }
    } else {
        white = b; k1 = Y;
        if (r <= g) { w1 = r - b; k2 = G; w2 = g - r; } else { w1 = g - b; k2 = R; w2 = r - g; }
    }
    auto mean = [](const float* v) {
        float a = 0.0f;
        for (int i = 0; i < 95; ++i) a += v[i];
        return a / 95.0f;
    };
    const float* p1 = crh::kRgbBasis[k1];
    const float* p2 = crh::kRgbBasis[k2];
    avg = white + w1 * mean(p1) + w2 * mean(p2);
    Spec out;
    for (int j = 0; j < 3; ++j) out.v[j] = white + w1 * p1[bins[j]] + w2 * p2[bins[j]];
    return out;
}
void upliftAll(float r, float g, float b, float* out95) {
    uint32_t bins[3];
    float avg;
    for (uint32_t i = 0; i < 95; ++i) {
        bins[0] = bins[1] = bins[2] = i;
        out95[i] = upliftBins(r, g, b, bins, avg).v[0];
    }
}

// ---------------------------------------------------------------------------
// Sampling warps (Math/Sampling.cu:4-58)
// ---------------------------------------------------------------------------
V2 sampleBarycentric(V2 u) {
    float u1 = u.x, u2 = u.y;
    bool flip = (u1 + u2) > 1.0f;
    u1 = flip ? (1.0f - u1) : u1;
    u2 = flip ? (1.0f - u2) : u2;
    return {u1, u2};
}
V3 cosineSampleHemisphere(V2 u) {  // :39-47
    float r = std::sqrt(u.x);
    float phi = 2.0f * kPi * u.y;
    float x = r * std::cos(phi);
    float y = r * std::sin(phi);
    float z = std::sqrt(fmax_std(0.0f, 1.0f - x * x - y * y));
    return {x, y, z};
}
float cosineHemispherePdf(V3 d) { return fmax_std(0.0f, d.z) * kInvPi; }
float powerHeuristic(int nf, float fPdf, int ng, float gPdf) {  // :53-58
    float f = nf * fPdf, g = ng * gPdf;
    return (f * f) / (f * f + g * g);
}
float acosSafe(float v) {  // Math/CRayMath.cu:11-14
    float lo = kEps, hi = float(1.0 - kEps);
    float c = v < lo ? lo : (hi < v ? hi : v);
    return std::acos(c);
}

// ---------------------------------------------------------------------------
// BSDFs (Material/*.cu)
// ---------------------------------------------------------------------------
struct BSDF {
    uint32_t type = CRH_MAT_DIFFUSE;
    Spec kd, ks;
    float n = 1.f;
    // Phong (PhongBRDF.cu:10-19)
    float pD = 0, pS = 0;
    // Principled (PrincipledBRDF.cu:47-58)
    float rough = 0, metal = 0;
    Frame frame;

    void init(const crh_material& m, const Frame& f) {  // BSDF.cu:6-32
        Spec d(m.diffuse[0], m.diffuse[1], m.diffuse[2]);
        Spec sp(m.specular[0], m.specular[1], m.specular[2]);
        initWith(m, f, d, sp, d.averageValue(), sp.averageValue());
    }
    // avgKd / avgKs: averageValue of the reflectances (all 95 bins in
    // spectral transport, where kd / ks hold the path's three bins).
    void initWith(const crh_material& m, const Frame& f, Spec d, Spec sp, float avgKd,
                  float avgKs) {
        frame = f;
        type = m.type;
        kd = d;
        ks = sp;
        n = m.shininess;
        if (type == CRH_MAT_PHONG) {
            pD = fmax_std(avgKd, 0.0f);
            pS = fmax_std(avgKs, 0.0f);
            float sum = fmax_std(pD + pS, kEps);
            pD /= sum;
            pS /= sum;
        } else if (type == CRH_MAT_PRINCIPLED) {
            float r = std::sqrt(2.0f / (n + 1.0f));
            float mt = avgKs;
            rough = std::clamp(r, 0.0f, 1.0f);
            metal = std::clamp(mt, 0.0f, 1.0f);
            float sp = 0.1f * (1.0f - metal) + 1.0f * metal;  // lerp(0.1,1,m)
            pD = 1.0f;
            pS = sp;
            pD /= (pD + pS);
            pS /= (pD + pS);  // uses the updated pD (reference behaviour)
        }
    }

    // --- Phong (PhongBRDF.cu:21-68)
    Spec phongEval(V3 wo, V3 wi) const {
        V3 r = reflect(-wo, V3(0, 0, 1));
        return kd * kInvPi + ks * (n + 2) * kInv2Pi *
                                 std::pow(fmax_std(dot(r, wi), 0.0f), n);
    }
    float phongPdf(V3 wo, V3 wi) const {
        float dp = cosineHemispherePdf(wi);
        V3 r = reflect(-wo, V3(0, 0, 1));
        float ca = fmax_std(dot(r, wi), 0.0f);
        float sp = (n + 1) * kInv2Pi * std::pow(ca, n);
        return dp * pD + sp * pS;
    }
    V3 phongSample(V3 wo, V2 u, float& pdf) const {
        V2 u0 = u;
        V3 wi(0.f);
        pdf = 0.0f;
        bool hasD = pD > 0.0f, hasS = pS > 0.0f;
        if (hasD && u0.x < pD) {
            u0.x /= pD;
            wi = cosineSampleHemisphere(u0);
            pdf = cosineHemispherePdf(wi);
        } else if (hasS) {
            u0.x = (u0.x - pD) / pS;
            float alpha = acosSafe(std::pow(u0.x, 1 / (n + 1)));
            float th = acosSafe(wo.z) + alpha;
            float hiPi = float(kPi / 2.0);
            th = th < 0.0f ? 0.0f : (hiPi < th ? hiPi : th);
            float phi = k2Pi * u0.y;
            float st = std::sin(th);
            wi = V3(st * std::cos(phi), st * std::sin(phi),
                    float(std::sqrt(1.0 - double(st * st))));
            pdf = (n + 1) * kInv2Pi * std::pow(std::cos(alpha), n);
        }
        return wi;
    }

    // --- Principled (PrincipledBRDF.cu:9-151)
    // ORC_MUTANT (0 = the restatement) plants one known mistake for the
    // tests that show the Principled closed forms catch it: 1 drops the
    // second 4*ci*cv division, 2 drops the Disney G remap, 3 un-inverts
    // the f0 lerp, 4 drops the 2 of fd90, 5 uses alpha = r (not r^2) in D.
    static float ggxD(float r, float ch) {  // :9-18
#if ORC_MUTANT == 5
        float a = r;
#else
        float a = r * r;
#endif
        float a2 = a * a;
        float c2 = ch * ch;
        float den = float(1.0 + (double(a2) - 1.0) * double(c2));
        return a2 * kInvPi / (den * den);
    }
    static float smithG1(float aG, float cv) {  // :20-24
        float a = aG * aG;
        float b = cv * cv;
        return 1 / (cv + std::sqrt(a + b - a * b));
    }
    static float smithG(float r, float ci, float cv) {  // :26-33
#if ORC_MUTANT == 2
        float a = r;
#else
        float a = float(0.5 + double(r) * 0.5);
#endif
        a *= a;
        return smithG1(a, ci) * smithG1(a, cv);
    }
    Spec prDiffuse(float cv, float ci, float cd) const {  // :135-141
#if ORC_MUTANT == 4
        float fd90 = float(0.5 + 1.0 * double(rough) * double(cd) * double(cd));
#else
        float fd90 = float(0.5 + 2.0 * double(rough) * double(cd) * double(cd));
#endif
        float fl = float(1.0 + (double(fd90) - 1.0) * std::pow(1.0 - double(ci), 5.0));
        float fv = float(1.0 + (double(fd90) - 1.0) * std::pow(1.0 - double(cv), 5.0));
        return kd * kInvPi * fv * fl;
    }
    Spec prSpecular(float cv, float ci, float ch, float cd) const {  // :143-150
        float D = ggxD(rough, ch);
#if ORC_MUTANT == 3
        Spec f0 = Spec(0.04f) * (1.0f - metal) + kd * metal;
#else
        Spec f0 = kd * (1.0f - metal) + Spec(0.04f) * metal;  // lerp(base, 0.04, m)
#endif
        float v0 = 1.f - cd;
        float v1 = v0 * v0;
        Spec F = f0 + (Spec(1.f) - f0) * v1 * v1 * v0;  // Fresnel.cu:3-8
        float G = smithG(rough, ci, cv);
#if ORC_MUTANT == 1
        float den = 1.0f;
#else
        float den = float(4.0 * double(ci) * double(cv));
#endif
        return D * F * G / den;
    }
    Spec prEval(V3 wo, V3 wi) const {  // :57-73
        if (wo.z <= 0.0 || wi.z <= 0.0) return Spec(0.f);
        V3 wh = normalize(wo + wi);
        float cv = wo.z, ci = wi.z, ch = wh.z;
        float cd = dot(wo, wh);
        Spec d = prDiffuse(cv, ci, cd) * (1.0f - metal);
        Spec sp = prSpecular(cv, ci, ch, cd);
        return d + sp;
    }
    float prPdf(V3 wo, V3 wi) const {  // :75-98
        if (wo.z <= 0.0 || wi.z <= 0.0) return 0.0f;
        V3 wh = normalize(wo + wi);
        float ch = wh.z;
        float cd = dot(wo, wh);
        if (cd <= 0.0f) return 0.0f;
        float dp = cosineHemispherePdf(wi);
        float sp = 0.0f;
        if (ch > 0.0f) {
            float D = ggxD(rough, ch);
            sp = D * ch / (4.0f * cd);
        }
        return dp * pD + sp * pS;
    }
    V3 prSample(V3 wo, V2 u, float& pdf) const {  // :100-133
        if (wo.z <= 0.0) {
            pdf = 0.0f;
            return V3(0.0f);
        }
        float u1 = u.x, u2 = u.y;
        V3 wi;
        if (u1 < pS) {
            u1 /= pS;
            float alpha = rough * rough;
            float a2 = alpha * alpha;
            float phi = u2 * k2Pi;
            float tan2 = a2 * u1 / (1 - u1);
            float ct = 1 / std::sqrt(1 + tan2);
            float r = std::sqrt(fmax_std(1 - ct * ct, 0.0f));
            V3 wh(std::cos(phi) * r, std::sin(phi) * r, ct);
            wi = reflect(-wo, wh);
        } else {
            u1 = (u1 - pS) / (1.0f - pS);
            wi = cosineSampleHemisphere(V2(u1, u2));
        }
        pdf = prPdf(wo, wi);
        return wi;
    }

    // --- dispatch in local space (BSDFBase, BSDF.h:15-58)
    Spec evalLocal(V3 wo, V3 wi) const {
        Spec f;
        if (type == CRH_MAT_DIFFUSE) f = kd * kInvPi;
        else if (type == CRH_MAT_PRINCIPLED) f = prEval(wo, wi);
        else f = phongEval(wo, wi);
        return f * std::fabs(wi.z);
    }
    float pdfLocal(V3 wo, V3 wi) const {
        if (type == CRH_MAT_DIFFUSE) return cosineHemispherePdf(wi);
        if (type == CRH_MAT_PRINCIPLED) return prPdf(wo, wi);
        return phongPdf(wo, wi);
    }
    V3 sampleLocal(V3 wo, V2 u, float& pdf) const {
        if (type == CRH_MAT_DIFFUSE) {
            V3 wi = cosineSampleHemisphere(u);
            pdf = cosineHemispherePdf(wi);
            return wi;
        }
        if (type == CRH_MAT_PRINCIPLED) return prSample(wo, u, pdf);
        return phongSample(wo, u, pdf);
    }

    // --- world-space wrappers (BSDF.cu:51-103)
    Spec evaluate(V3 wo, V3 wi) const {
        return evalLocal(frame.toLocal(wo), frame.toLocal(wi));
    }
    float evaluatePdf(V3 wo, V3 wi) const {
        return pdfLocal(frame.toLocal(wo), frame.toLocal(wi));
    }
    Spec sampleEvaluate(Sampler& smp, V3 wo, V3& wi, float& pdf) const {
        V3 woL = frame.toLocal(wo);
        pdf = 0.0f;
        V2 u = smp.next2D();
        V3 wiL = sampleLocal(woL, u, pdf);
        Spec f = evalLocal(woL, wiL);
        wi = frame.toWorld(wiL);
        return f;
    }
};

// ---------------------------------------------------------------------------
// Guided continuation (an addition with no reference counterpart): one-sample
// MIS of the reference's BSDF sampling (fraction alpha) and a proposal, the
// guide's 8-lobe von Mises-Fisher mixture, as the product's guided path does
// it (craystal_amd/csrc/wavefront.hip kWfShadeCont, guide_device.h).
// Pinned by: the actual density of each BSDF sampling procedure against
// histograms of its samples; the mixture against the float64 model of
// oracle/guiding.py (itself pinned by quadrature and histograms); a guide
// equal to the BSDF's own procedure giving the plain estimator's weight for
// every direction; guided images against closed forms and plain images.
// ---------------------------------------------------------------------------
struct Mixture {
    float w[8];
    V3 mu[8];
    float k[8];
};

// The network's head from its 48 outputs (5 per lobe: weight logit, mean,
// kappa): softmax weights, unit means (length floored at 1e-6), kappa =
// softplus + 0.01.
Mixture mixtureFromOutputs(const float* o) {
    Mixture m;
    float amax = -INFINITY;
    for (int j = 0; j < 8; ++j) amax = std::max(amax, o[5 * j]);
    float sum = 0.f;
    for (int j = 0; j < 8; ++j) {
        m.w[j] = std::exp(o[5 * j] - amax);
        sum += m.w[j];
    }
    for (int j = 0; j < 8; ++j) {
        m.w[j] /= sum;
        float mx = o[5 * j + 1], my = o[5 * j + 2], mz = o[5 * j + 3];
        float len = std::max(std::sqrt(mx * mx + my * my + mz * mz), 1e-6f);
        m.mu[j] = V3(mx / len, my / len, mz / len);
        float x = o[5 * j + 4];
        m.k[j] = (x > 20.f ? x : std::log1p(std::exp(std::min(x, 20.f)))) + 0.01f;
    }
    return m;
}

// vMF density k / (2 pi (1 - e^{-2k})) exp(k (mu.d - 1)) (stable form).
float vmfDensity(float k, float c) {
    return k / (2.f * kPi * (-std::expm1(-2.f * k))) * std::exp(k * (c - 1.f));
}
float mixturePdf(const Mixture& m, V3 d) {
    float q = 0.f;
    for (int j = 0; j < 8; ++j) q += m.w[j] * vmfDensity(m.k[j], dot(m.mu[j], d));
    return q;
}
// Lobe by u0 against the weights' CDF (the last lobe if rounding leaves u0
// above it); direction by W. Jakob's stable inversion (2012) from (u1, u2).
V3 mixtureSample(const Mixture& m, float u0, float u1, float u2) {
    int j = 7;
    float c = 0.f;
    for (int i = 0; i < 8; ++i) {
        c += m.w[i];
        if (u0 < c) {
            j = i;
            break;
        }
    }
    float k = m.k[j];
    V3 mu = m.mu[j];
    float w = 1.f + std::log(u1 + (1.f - u1) * std::exp(-2.f * k)) / k;
    w = std::min(std::max(w, -1.f), 1.f);
    float sn = std::sqrt(std::max(0.f, 1.f - w * w));
    V3 a = std::fabs(mu.x) > 0.9f ? V3(0, 1, 0) : V3(1, 0, 0);
    V3 t = normalize(cross(a, mu));
    V3 b = cross(mu, t);
    float phi = 2.f * kPi * u2;
    return normalize(mu * w + (t * std::cos(phi) + b * std::sin(phi)) * sn);
}

// Density of the directions BSDF::sampleLocal produces (which the reference
// reports only for Lambertian): Principled picks the GGX half-vector lobe
// with probability pS (PrincipledBRDF.cu:112), whose reflected direction has
// density D(h) cos(theta_h) / (4 |wo.h|), else the cosine lobe.
float actualPdfLocal(const BSDF& b, V3 wo, V3 wi) {
    if (b.type == CRH_MAT_DIFFUSE) return cosineHemispherePdf(wi);
    if (b.type == CRH_MAT_PRINCIPLED) {
        if (wo.z <= 0.0f) return 0.f;
        float spec = 0.f;
        // The half vector the sampler drew is the one of +-(wo + wi) with
        // positive z (the other sign maps to the same wi).
        V3 h = normalize(wo + wi);
        if (h.z < 0.f) h = -h;
        float cd = dot(wo, h);
        if (h.z > 0.f && std::fabs(cd) > 0.f)
            spec = BSDF::ggxD(b.rough, h.z) * h.z / (4.f * std::fabs(cd));
        return b.pS * spec + (1.f - b.pS) * cosineHemispherePdf(wi);
    }
    return 0.f;  // Phong is never guided
}

struct Guide {
    Mixture mix;
    float alpha = 0.5f;     // probability of the BSDF's own sample
    uint32_t mask = 3u;     // bit t: guide BSDF type t (0 Lambertian, 1 Principled)
    uint32_t maxDepth = 0;  // guide vertices at depth < maxDepth (0: all)
    bool bsdfAsGuide = false;  // the proposal is the BSDF's own procedure
    bool applies(const BSDF& b, uint32_t depth) const {
        return b.type != CRH_MAT_PHONG && ((mask >> b.type) & 1u) &&
               (maxDepth == 0u || depth < maxDepth);
    }
};

// One guided continuation.  Draws, in order: the technique (1), then the
// BSDF's 2D sample, or the lobe (1) and the direction (2) of the proposal
// (the BSDF-as-guide proposal draws the same three and samples the BSDF
// from the last two).  Returns f * |cos| at wiL (local; wi is the world
// direction) and sets the pdf whose ratio f / pdf is the sample's weight:
// q_mix * p_rep / p_act with q_mix = alpha p_act + (1 - alpha) q_guide, so
// E[f / pdf] equals the plain estimator's E[f / p_rep] over wi ~ p_act.
template <class S>
Spec guidedSample(const BSDF& b, const Guide& g, S& smp, V3 wo, V3& wi, float& pdf) {
    V3 woL = b.frame.toLocal(wo), wiL;
    float dummy = 0.f;
    if (smp.next1D() < g.alpha) {
        V2 u = smp.next2D();
        wiL = b.sampleLocal(woL, u, dummy);
        wi = b.frame.toWorld(wiL);
    } else {
        float u0 = smp.next1D();
        V2 u = smp.next2D();
        if (g.bsdfAsGuide) {
            (void)u0;
            wiL = b.sampleLocal(woL, u, dummy);
            wi = b.frame.toWorld(wiL);
        } else {
            wi = mixtureSample(g.mix, u0, u.x, u.y);
            wiL = b.frame.toLocal(wi);
        }
    }
    Spec f = b.evalLocal(woL, wiL);
    float prep = b.pdfLocal(woL, wiL);
    float pact = actualPdfLocal(b, woL, wiL);
    float q = g.bsdfAsGuide ? pact : mixturePdf(g.mix, wi);
    float qmix = g.alpha * pact + (1.f - g.alpha) * q;
    pdf = (prep > 0.f && pact > 0.f && qmix > 0.f) ? qmix * prep / pact : 0.f;
    return f;
}

// Uniforms given by the caller (for the unit pins), in the order drawn.
struct FixedUniforms {
    const float* u;
    int i = 0;
    float next1D() { return u[i++]; }
    V2 next2D() {
        V2 r(u[i], u[i + 1]);
        i += 2;
        return r;
    }
};

// ---------------------------------------------------------------------------
// Camera (Core/Camera.cu:35-39, Core/CameraProxy.h:25-40)
// ---------------------------------------------------------------------------
void cameraUpdate(crh_camera& c) {
    V3 pos(c.pos[0], c.pos[1], c.pos[2]);
    V3 tgt(c.target[0], c.target[1], c.target[2]);
    V3 up(c.up[0], c.up[1], c.up[2]);
    V3 W = normalize(tgt - pos);
    V3 U = normalize(cross(W, up));
    V3 V = normalize(cross(U, W));
    for (int i = 0; i < 3; ++i) {
        c.w[i] = W[i];
        c.u[i] = U[i];
        c.v[i] = V[i];
    }
}

Ray generateRay(const crh_camera& c, uint32_t W, uint32_t H, V2 sp) {
    float u = (2.0f * (sp.x / float(W)) - 1.0f);
    float v = (1.0f - 2.0f * (sp.y / float(H)));
    float aspect = float(W) / float(H);
    // fovY * 0.5 and glm::radians are double; tan in double.
    float hh = float(std::tan(double(c.fov_y) * 0.5 *
                              0.01745329251994329576923690768489));
    float hw = aspect * hh;
    V3 U(c.u[0], c.u[1], c.u[2]), Vv(c.v[0], c.v[1], c.v[2]),
        Wv(c.w[0], c.w[1], c.w[2]);
    V3 dir = normalize(U * (u * hw) + Vv * (v * hh) + Wv);
    Ray r;
    r.o = V3(c.pos[0], c.pos[1], c.pos[2]);
    r.d = dir;
    return r;
}

// ---------------------------------------------------------------------------
// Integrator (Core/Integrator.cu:13-227)
// ---------------------------------------------------------------------------
struct LightSample {
    Spec weight;
    V3 dirW;
    float pdf = 0;
    uint32_t emissiveID = CRH_INVALID_PRIM;
    float lightSelectPdf = 0.0f;
};

struct Tally {
    uint64_t rays = 0;
    // MIS BSDF rays that cannot have the sampled emitter triangle as their
    // closest hit (its Moller-Trumbore test fails, no spheres): a renderer
    // may skip tracing them, since evalMIS adds nothing for them.  Counted
    // only; every ray is still traced here.
    uint64_t misUnreachable = 0;
};

inline const crh_material& matOf(const Scene& s, uint32_t prim) {
    return s.mats[s.meshMat[s.primMesh[prim]]];
}
inline uint32_t matIdOf(const Scene& s, uint32_t prim) { return s.meshMat[s.primMesh[prim]]; }
// The material of a hit: getMeshDesc(primitiveIndex) whatever the hit type
// (Core/Integrator.cu:172-178), unless the scene gives spheres materials.
inline uint32_t matIdOfHit(const Scene& s, const RayHit& rh) {
    if (rh.sphere && !s.sphereMat.empty()) return s.sphereMat[rh.prim];
    return matIdOf(s, rh.prim);
}

// Spectral transport: a path carries the bins h, h + 32, h + 64 (mod 95) of
// its hero bin h in the three Spec channels.
struct Wl {
    bool on = false;
    uint32_t b[3] = {0, 0, 0};
};
inline Spec lanesOf(const float* p, const Wl& w) { return Spec(p[w.b[0]], p[w.b[1]], p[w.b[2]]); }
inline Spec emissionOfMat(const Scene& s, uint32_t mid, const Wl& w) {
    if (w.on) return lanesOf(s.spectra.data() + (size_t(mid) * 3 + 2) * 95, w);
    const crh_material& m = s.mats[mid];
    return Spec(m.emission[0], m.emission[1], m.emission[2]);
}
inline Spec emissionOf(const Scene& s, uint32_t prim, const Wl& w) {
    return emissionOfMat(s, matIdOf(s, prim), w);
}

// SceneView::intersect as the integrator uses it.  A reported hit without a
// recorded primitive (all-NaN comparisons) counts as a miss; the reference
// would read an uninitialised primitive index (Scene/Scene.cu:31-58).
bool intersect(const Scene& s, RayHit& rh, Tally& tl) {
    tl.rays++;
    return sceneIntersect(s, rh, nullptr, false) && rh.prim != CRH_INVALID_PRIM;
}
bool occluded(const Scene& s, const Ray& r, Tally& tl) {
    tl.rays++;
    RayHit rh;
    rh.ray = r;
    return sceneIntersect(s, rh, nullptr, true);
}

LightSample sampleLight(const Scene& s, const Isect& it, const BSDF& bsdf,
                        Sampler& smp, Tally& tl, const Wl& wl) {  // :21-72
    int cnt = int(s.emissive.size());
    if (cnt == 0) return {};
    int sel = std::min<int>(int(float(cnt) * smp.next1D()), cnt - 1);
    uint32_t eid = s.emissive[sel];
    Tri tri = getTriangle(s, eid);
    V2 b = sampleBarycentric(smp.next2D());
    Vtx vd = tri.interpolate(V3(1.f - b.x - b.y, b.x, b.y));
    V3 target = offsetRayOrigin(vd.p, tri.faceNormal());
    V3 toLight = target - it.posW;
    float d2 = dot(toLight, toLight);
    float dist = std::sqrt(d2);
    V3 ld = toLight / dist;
    if (dot(it.orientedFaceNormal(), ld) < 0.0 || dot(vd.n, ld) > 0.0 ||
        dist < kEps)
        return {};
    Ray sr;
    sr.o = offsetRayOrigin(it.posW, it.orientedFaceNormal());
    sr.d = ld;
    sr.tMax = length(sr.o - target) - kEps;
    if (occluded(s, sr, tl)) return {};
    LightSample ls;
    ls.weight = bsdf.evaluate(it.viewW, ld) * emissionOf(s, eid, wl);
    ls.dirW = ld;
    ls.pdf = d2 / (tri.area() * absDot(ld, vd.n));
    ls.emissiveID = eid;
    ls.lightSelectPdf = float(1.0 / cnt);
    return ls;
}

Spec evalMIS(const Scene& s, const Isect& it, const BSDF& bsdf, Sampler& smp,
             Tally& tl, const Wl& wl) {  // :74-140
    Spec value(0.0f);
    LightSample ls = sampleLight(s, it, bsdf, smp, tl, wl);
    if (ls.pdf <= 0.0) return value;
    {
        float lightPdf = ls.pdf;
        if (lightPdf > 0.0) {
            float bsdfPdf = bsdf.evaluatePdf(it.viewW, ls.dirW);
            if (bsdfPdf > 0.0)
                value = value + powerHeuristic(1, lightPdf, 1, bsdfPdf) *
                                    ls.weight / lightPdf;
        }
    }
    {
        V3 dir;
        float bsdfPdf = 0.0f;
        Spec bw = bsdf.sampleEvaluate(smp, it.viewW, dir, bsdfPdf);
        if (bw.maxValue() > 0.0 && bsdfPdf > 0.0) {
            RayHit rh;
            rh.ray.o = offsetRayOrigin(it.posW, it.orientedFaceNormal());
            rh.ray.d = dir;
            if (s.sphC.empty()) {
                float eu = 0, ev = 0, et = kFltInf;
                if (!(triIntersect(s, ls.emissiveID, rh.ray, eu, ev, et) &&
                      !(et < rh.ray.tMin || et > rh.ray.tMax) && et < kFltInf))
                    tl.misUnreachable++;
            }
            // :108 compares primitive indices only: with the reference's
            // sphere materials a sphere of the light's index matches too.
            if (intersect(s, rh, tl) && rh.prim == ls.emissiveID &&
                (!rh.sphere || s.sphereMat.empty())) {
                Isect lit = createIntersection(s, rh);
                const uint32_t mid = matIdOfHit(s, rh);
                const crh_material& m = s.mats[mid];
                if (m.flags & CRH_MATF_EMISSIVE) {
                    Spec em = emissionOfMat(s, mid, wl);
                    V3 toLight = lit.posW - it.posW;
                    float d2 = dot(toLight, toLight);
                    V3 lv = toLight / std::sqrt(d2);
                    float cl = absDot(lv, lit.faceN);
                    if (cl != 0.0) {
                        float area = getTriangle(s, rh.prim).area();
                        float lightPdf = d2 / (area * cl);
                        value = value + powerHeuristic(1, bsdfPdf, 1, lightPdf) *
                                            em * bw / bsdfPdf;
                    }
                }
            }
        }
    }
    return value / ls.lightSelectPdf;
}

// pathTraceKernel body (:142-227) for one (pixel, sample index).
// guide (optional): the guided continuation at the vertices it applies to.
Spec tracePath(const Scene& s, const crh_camera& cam, uint32_t W, uint32_t H,
               const crh_integrator_cfg& cfg, uint32_t x, uint32_t y,
               uint32_t sampleIndex, V2& pixelOut, Tally& tl, const Guide* guide = nullptr) {
    Sampler smp(x, y, sampleIndex);
    V2 j = smp.next2D();
    V2 pixel(float(x) + j.x, float(y) + j.y);
    pixelOut = pixel;
    Wl wl;
    if (cfg.spectral) {  // hero bin drawn right after the jitter
        uint32_t h = std::min<uint32_t>(uint32_t(smp.next1D() * 95.0f), 94u);
        wl.on = true;
        wl.b[0] = h;
        wl.b[1] = (h + 32u) % 95u;
        wl.b[2] = (h + 64u) % 95u;
    }
    Ray ray = generateRay(cam, W, H, pixel);
    Spec radiance(0.f), beta(1.f);
    const float rr = cfg.rr_threshold;
    for (uint32_t depth = 0; depth < cfg.max_depth; ++depth) {
        RayHit rh;
        rh.ray = ray;
        bool terminate = false;
        if (intersect(s, rh, tl)) {
            Isect it = createIntersection(s, rh);
            const uint32_t mid = matIdOfHit(s, rh);
            const crh_material& m = s.mats[mid];
            if (m.flags & CRH_MATF_EMISSIVE) {
                radiance = radiance + emissionOfMat(s, mid, wl) * beta;
                break;
            }
            BSDF bsdf;  // getBSDF (Material/BSDF.cu:6-32)
            const float* sp = wl.on ? s.spectra.data() + size_t(mid) * 3 * 95 : nullptr;
            if (m.flags & CRH_MATF_DIFFUSE_TEXTURE) {
                F4 t = textureSample(s.textures[m.diffuse_tex], it.uv);
                if (wl.on) {
                    float avgKd;
                    Spec kd = upliftBins(t.v[0], t.v[1], t.v[2], wl.b, avgKd);
                    bsdf.initWith(m, it.frame, kd, lanesOf(sp + 95, wl), avgKd, s.avgKs[mid]);
                } else {
                    Spec kd(t.v[0], t.v[1], t.v[2]);
                    Spec ks(m.specular[0], m.specular[1], m.specular[2]);
                    bsdf.initWith(m, it.frame, kd, ks, kd.averageValue(), ks.averageValue());
                }
            } else if (wl.on) {
                bsdf.initWith(m, it.frame, lanesOf(sp, wl), lanesOf(sp + 95, wl), s.avgKd[mid],
                              s.avgKs[mid]);
            } else {
                bsdf.init(m, it.frame);
            }
            if (cfg.use_mis) {
                radiance = radiance + evalMIS(s, it, bsdf, smp, tl, wl) * beta;
            } else {
                LightSample ls = sampleLight(s, it, bsdf, smp, tl, wl);
                if (ls.pdf > 0.0)
                    radiance = radiance +
                               ls.weight * beta / ls.pdf / ls.lightSelectPdf;
            }
            V3 wi;
            float pdf = 0.0f;
            Spec f = guide && guide->applies(bsdf, depth)
                         ? guidedSample(bsdf, *guide, smp, it.viewW, wi, pdf)
                         : bsdf.sampleEvaluate(smp, it.viewW, wi, pdf);
            if (pdf <= 0.0f) break;
            beta = beta * (f / pdf);
            ray.o = offsetRayOrigin(it.posW, it.orientedFaceNormal());
            ray.d = wi;
        } else {
            // Constant environment (an addition; zero = the reference, which returns 0).
            radiance = radiance + (wl.on ? lanesOf(s.envSpec.data(), wl) : s.env) * beta;
            terminate = true;
        }
        if (smp.next1D() < rr) terminate = true;
        if (terminate) break;
        if (depth >= cfg.rr_start_depth) {  // reference: depth > 5 (start 6)
            float q = fmax_std(rr, 1.0f - beta.maxValue());
            q = std::clamp(q, 0.0f, 1.0f);
            if (smp.next1D() < q) break;
            beta = beta / (1.0f - q);
        }
    }
    if (wl.on) {  // CIE XYZ of the three lanes
        const float* a = s.filmW[wl.b[0]];
        const float* b = s.filmW[wl.b[1]];
        const float* c = s.filmW[wl.b[2]];
        const float* L = radiance.v;
        radiance = Spec((L[0] * a[0] + L[1] * b[0]) + L[2] * c[0],
                        (L[0] * a[1] + L[1] * b[1]) + L[2] * c[1],
                        (L[0] * a[2] + L[1] * b[2]) + L[2] * c[2]);
    }
    return radiance;
}

// SensorData::addSample (Core/Sensor.cu:19-43), one sample, fixed order.
void splat(float* sum, float* wsum, uint32_t W, uint32_t H, const Spec& L,
           V2 xy) {
    if (L.hasNaN() || L.hasInf()) return;
    int x0 = int(std::floor(xy.x - 1.0f)), y0 = int(std::floor(xy.y - 1.0f));
    int x1 = int(std::ceil(xy.x + 1.0f)), y1 = int(std::ceil(xy.y + 1.0f));
    x0 = std::max(x0, 0);
    y0 = std::max(y0, 0);
    x1 = std::min(x1, int(W) - 1);
    y1 = std::min(y1, int(H) - 1);
    for (int y = y0; y <= y1; ++y)
        for (int x = x0; x <= x1; ++x) {
            float dx = xy.x - (float(x) + 0.5f);
            float dy = xy.y - (float(y) + 0.5f);
            float w = fmax_std(0.0f, 1.0f - std::fabs(dx)) *
                      fmax_std(0.0f, 1.0f - std::fabs(dy));
            size_t id = size_t(W) * y + x;
            for (int c = 0; c < 3; ++c) sum[3 * id + c] += L.v[c] * w;
            wsum[id] += w;
        }
}

Scene* makeScene(const crh_scene_desc& d) {
    Scene* s = new Scene();
    for (uint32_t m = 0; m < d.nmesh; ++m) {
        const crh_mesh& mesh = d.meshes[m];
        uint32_t voff = uint32_t(s->pos.size());
        for (uint32_t i = 0; i < mesh.nvert; ++i) {
            s->pos.push_back(V3(mesh.position[3 * i], mesh.position[3 * i + 1],
                                mesh.position[3 * i + 2]));
            s->nrm.push_back(V3(mesh.normal[3 * i], mesh.normal[3 * i + 1],
                                mesh.normal[3 * i + 2]));
            s->uv.push_back(mesh.texcrd ? V2(mesh.texcrd[2 * i], mesh.texcrd[2 * i + 1])
                                        : V2(0.f, 0.f));
        }
        for (uint32_t i = 0; i < mesh.nindex; ++i)
            s->idx.push_back(mesh.index[i] + voff);
        for (uint32_t i = 0; i < mesh.nindex / 3; ++i) s->primMesh.push_back(m);
        s->meshMat.push_back(mesh.material_id);
    }
    s->mats.assign(d.materials, d.materials + d.nmaterial);
    s->emissive.assign(d.emissive_index, d.emissive_index + d.nemissive);
    s->env = Spec(d.env_radiance[0], d.env_radiance[1], d.env_radiance[2]);
    for (uint32_t i = 0; i < d.nsphere; ++i) {
        s->sphC.push_back(V3(d.sphere_center[3 * i], d.sphere_center[3 * i + 1],
                             d.sphere_center[3 * i + 2]));
        s->sphR.push_back(d.sphere_radius[i]);
    }
    if (d.nsphere && d.sphere_material)
        s->sphereMat.assign(d.sphere_material, d.sphere_material + d.nsphere);
    for (uint32_t t = 0; t < d.ntexture; ++t) {  // Texture(const Image&), Core/Texture.cu:46-62
        const crh_texture& tx = d.textures[t];
        Scene::Tex o{tx.width, tx.height, std::vector<float>(4 * size_t(tx.width) * tx.height, 0.f)};
        for (size_t i = 0; i < size_t(tx.width) * tx.height; ++i)
            for (uint32_t c = 0; c < tx.channels; ++c) o.px[4 * i + c] = tx.data[i * tx.channels + c];
        s->textures.push_back(std::move(o));
    }
    // Spectral transport: the given spectra, or every RGB value uplifted.
    s->spectral = true;
    if (d.spectra) {
        s->spectra.assign(d.spectra, d.spectra + size_t(d.nmaterial) * 3 * 95);
    } else {
        s->spectra.assign(size_t(d.nmaterial) * 3 * 95, 0.0f);
        for (uint32_t m = 0; m < d.nmaterial; ++m) {
            const crh_material& mt = d.materials[m];
            float* o = s->spectra.data() + size_t(m) * 285;
            upliftAll(mt.diffuse[0], mt.diffuse[1], mt.diffuse[2], o);
            upliftAll(mt.specular[0], mt.specular[1], mt.specular[2], o + 95);
            upliftAll(mt.emission[0], mt.emission[1], mt.emission[2], o + 190);
        }
    }
    {
        s->envSpec.assign(95, 0.0f);
        if (d.spectra && d.env_spectrum) s->envSpec.assign(d.env_spectrum, d.env_spectrum + 95);
        if (!d.spectra) upliftAll(d.env_radiance[0], d.env_radiance[1], d.env_radiance[2],
                                  s->envSpec.data());
        for (uint32_t m = 0; m < d.nmaterial; ++m) {
            float a = 0.0f, b = 0.0f;
            for (int i = 0; i < 95; ++i) a += s->spectra[size_t(m) * 285 + i];
            for (int i = 0; i < 95; ++i) b += s->spectra[size_t(m) * 285 + 95 + i];
            s->avgKd.push_back(a / 95.0f);
            s->avgKs.push_back(b / 95.0f);
        }
        // Hero-bin film weight of one of a path's three bins: the XYZ sum's
        // factor (kCIENormalizationY * kXYZStep) times 95 / 3.
        const double k = double(crh::kCieNormY * crh::kCieXYZStep) * 95.0 / 3.0;
        double c[3];
        for (int i = 0; i < 95; ++i) {
            cieXYZ(i, c);
            for (int j = 0; j < 3; ++j) s->filmW[i][j] = float(c[j] * k);
        }
    }
    s->bvh = buildBVH(s->pos, s->idx, d.bvh.bins, d.bvh.max_leaf,
                      d.bvh.max_depth);
    return s;
}

void fillStats(const BVHOut& b, crh_bvh_stats* st) {
    std::memset(st, 0, sizeof(*st));
    st->node_count = uint32_t(b.nodes.size());
    st->max_depth = b.maxDepth;
    st->max_leaf = b.maxLeaf;
    st->root = b.root;
    st->prim_count = uint32_t(b.primOrder.size());
    uint32_t ni = 0;
    for (auto& n : b.nodes) ni += n.tag == 1;
    st->internal_count = ni;
    for (int i = 0; i < 3; ++i) {
        st->root_min[i] = b.rootBounds.pMin[i];
        st->root_max[i] = b.rootBounds.pMax[i];
    }
    st->build_ms = b.buildMs;
}

void exportNodes(const BVHOut& b, crh_bvh_node_ref* nodes, uint32_t* prims) {
    if (nodes)
        for (size_t i = 0; i < b.nodes.size(); ++i) {
            nodes[i].tag = b.nodes[i].tag;
            nodes[i].a = b.nodes[i].a;
            nodes[i].b = b.nodes[i].b;
            for (int k = 0; k < 3; ++k) {
                nodes[i].pmin[k] = b.nodes[i].bounds.pMin[k];
                nodes[i].pmax[k] = b.nodes[i].bounds.pMax[k];
            }
        }
    if (prims) std::copy(b.primOrder.begin(), b.primOrder.end(), prims);
}

inline Ray toRay(const crh_ray& r) {
    Ray o;
    o.o = V3(r.origin[0], r.origin[1], r.origin[2]);
    o.tMin = r.tmin;
    o.d = V3(r.dir[0], r.dir[1], r.dir[2]);
    o.tMax = r.tmax;
    return o;
}

}  // namespace orc

using namespace orc;

extern "C" {

// ---- primitives -----------------------------------------------------------
void orc_sampler_seq(uint32_t x, uint32_t y, uint32_t idx, uint32_t n,
                     float* out, uint32_t* raw) {
    Sampler s(x, y, idx);
    Sampler s2(x, y, idx);
    for (uint32_t i = 0; i < n; ++i) {
        if (out) out[i] = s.next1D();
        if (raw) raw[i] = s2.next();
    }
}

void orc_camera_update(crh_camera* c) { cameraUpdate(*c); }

void orc_camera_ray(const crh_camera* c, uint32_t W, uint32_t H, float px,
                    float py, crh_ray* out) {
    Ray r = generateRay(*c, W, H, V2(px, py));
    for (int i = 0; i < 3; ++i) {
        out->origin[i] = r.o[i];
        out->dir[i] = r.d[i];
    }
    out->tmin = r.tMin;
    out->tmax = r.tMax;
}

void orc_offset_ray_origin(const float* p, const float* n, float* out) {
    V3 r = offsetRayOrigin(V3(p[0], p[1], p[2]), V3(n[0], n[1], n[2]));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

int orc_aabb_intersect(const float* pmin, const float* pmax, const crh_ray* r,
                       float* t) {
    AABB b;
    b.pMin = V3(pmin[0], pmin[1], pmin[2]);
    b.pMax = V3(pmax[0], pmax[1], pmax[2]);
    float ht = kFltInf;
    bool h = aabbIntersect(b, toRay(*r), ht);
    *t = ht;
    return h;
}

// ---- BVH ------------------------------------------------------------------
void* orc_bvh_build(const float* pos, uint32_t nvert, const uint32_t* idx,
                    uint32_t ntri, uint32_t bins, uint32_t maxLeaf,
                    uint32_t maxDepth) {
    std::vector<V3> p(nvert);
    for (uint32_t i = 0; i < nvert; ++i)
        p[i] = V3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    std::vector<uint32_t> id(idx, idx + size_t(ntri) * 3);
    return new BVHOut(buildBVH(p, id, bins, maxLeaf, maxDepth));
}
void orc_bvh_stats(void* h, crh_bvh_stats* st) { fillStats(*(BVHOut*)h, st); }
void orc_bvh_export(void* h, crh_bvh_node_ref* nodes, uint32_t* prims) {
    exportNodes(*(BVHOut*)h, nodes, prims);
}
void orc_bvh_free(void* h) { delete (BVHOut*)h; }

// ---- scene ----------------------------------------------------------------
void* orc_scene_create(const crh_scene_desc* d) { return makeScene(*d); }
void orc_scene_free(void* s) { delete (Scene*)s; }
void orc_scene_bvh_stats(void* s, crh_bvh_stats* st) {
    fillStats(((Scene*)s)->bvh, st);
}
void orc_scene_bvh_export(void* s, crh_bvh_node_ref* nodes, uint32_t* prims) {
    exportNodes(((Scene*)s)->bvh, nodes, prims);
}

// Closest hit per ray; node_visits / tri_tests (optional) per ray.
// crh_hit of a RayHit: sphere hits carry CRH_HIT_SPHERE and u = v = 0.
static void writeHit(const RayHit& rh, crh_hit& h) {
    const bool none = rh.prim == CRH_INVALID_PRIM;
    h.prim = none ? rh.prim : (rh.prim | (rh.sphere ? CRH_HIT_SPHERE : 0u));
    h.t = rh.hitT;
    h.u = none || rh.sphere ? 0.f : rh.bary.x;
    h.v = none || rh.sphere ? 0.f : rh.bary.y;
}

void orc_trace_closest(void* sp, const crh_ray* rays, uint64_t n,
                       crh_hit* hits, uint32_t* node_visits,
                       uint32_t* tri_tests, int nthreads) {
    const Scene& s = *(Scene*)sp;
    // The thread count applies to this call only (num_threads clause), not to
    // later calls.
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nt)
    for (int64_t i = 0; i < int64_t(n); ++i) {
        RayHit rh;
        rh.ray = toRay(rays[i]);
        Counters c;
        sceneIntersect(s, rh, &c, false);
        writeHit(rh, hits[i]);
        if (node_visits) node_visits[i] = uint32_t(c.nodes);
        if (tri_tests) tri_tests[i] = uint32_t(c.tris);
    }
}

void orc_trace_any(void* sp, const crh_ray* rays, uint64_t n, uint8_t* occ,
                   int nthreads) {
    const Scene& s = *(Scene*)sp;
    // The thread count applies to this call only (num_threads clause), not to
    // later calls.
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nt)
    for (int64_t i = 0; i < int64_t(n); ++i) {
        RayHit rh;
        rh.ray = toRay(rays[i]);
        occ[i] = sceneIntersect(s, rh, nullptr, true) ? 1 : 0;
    }
}

// Brute force over every triangle in primitive-ID order (the pin for the
// traversal: same t; same ID except on exact t ties).
void orc_trace_brute(void* sp, const crh_ray* rays, uint64_t n, crh_hit* hits) {
    const Scene& s = *(Scene*)sp;
    uint32_t ntri = uint32_t(s.idx.size() / 3);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < int64_t(n); ++i) {
        RayHit rh;
        rh.ray = toRay(rays[i]);
        for (uint32_t k = 0; k < uint32_t(s.sphR.size()); ++k) intersectSphereShape(s, k, rh);
        for (uint32_t t = 0; t < ntri; ++t) intersectShape(s, t, rh);
        writeHit(rh, hits[i]);
    }
}

// ---- BSDF in local space (golden vectors) -----------------------------------
// out_f[3] = evaluate (incl. |cos|), *out_pdf = evaluatePdf.
void orc_bsdf_eval_local(const crh_material* m, const float* wo,
                         const float* wi, float* out_f, float* out_pdf) {
    BSDF b;
    b.init(*m, Frame());
    V3 o(wo[0], wo[1], wo[2]), i(wi[0], wi[1], wi[2]);
    Spec f = b.evalLocal(o, i);
    for (int k = 0; k < 3; ++k) out_f[k] = f.v[k];
    *out_pdf = b.pdfLocal(o, i);
}
// Sample with explicit u; returns wi, f (evaluate at wi) and the sample pdf.
void orc_bsdf_sample_local(const crh_material* m, const float* wo,
                           const float* u, float* out_wi, float* out_f,
                           float* out_pdf) {
    BSDF b;
    b.init(*m, Frame());
    V3 o(wo[0], wo[1], wo[2]);
    float pdf = 0.f;
    V3 wi = b.sampleLocal(o, V2(u[0], u[1]), pdf);
    Spec f = b.evalLocal(o, wi);
    for (int k = 0; k < 3; ++k) {
        out_wi[k] = wi[k];
        out_f[k] = f.v[k];
    }
    *out_pdf = pdf;
}

void orc_film_splat(float* sum, float* wsum, uint32_t W, uint32_t H,
                    const float* L, float px, float py) {
    splat(sum, wsum, W, H, Spec(L[0], L[1], L[2]), V2(px, py));
}

// Per-sample radiance of one pass (sample index `si`) for every pixel.
// L: W*H*3 floats.  Returns rays traced.
static uint64_t renderSamples(void* sp, const crh_camera* cam, uint32_t W, uint32_t H,
                              const crh_integrator_cfg* cfg, uint32_t si, float* L, float* pix,
                              int nthreads, const Guide* guide);

static uint64_t renderSamplesCounted(void* sp, const crh_camera* cam, uint32_t W, uint32_t H,
                                     const crh_integrator_cfg* cfg, uint32_t si, float* L,
                                     float* pix, int nthreads, const Guide* guide,
                                     uint64_t* misUnreachable);

uint64_t orc_render_samples(void* sp, const crh_camera* cam, uint32_t W,
                            uint32_t H, const crh_integrator_cfg* cfg,
                            uint32_t si, float* L, float* pix, int nthreads) {
    return renderSamples(sp, cam, W, H, cfg, si, L, pix, nthreads, nullptr);
}

// As orc_render_samples; also returns in *misUnreachable the MIS BSDF rays
// whose sampled emitter triangle they cannot hit (see Tally).
uint64_t orc_render_samples_counted(void* sp, const crh_camera* cam, uint32_t W, uint32_t H,
                                    const crh_integrator_cfg* cfg, uint32_t si, float* L,
                                    float* pix, int nthreads, uint64_t* misUnreachable) {
    return renderSamplesCounted(sp, cam, W, H, cfg, si, L, pix, nthreads, nullptr,
                                misUnreachable);
}

static uint64_t renderSamples(void* sp, const crh_camera* cam, uint32_t W, uint32_t H,
                              const crh_integrator_cfg* cfg, uint32_t si, float* L, float* pix,
                              int nthreads, const Guide* guide) {
    return renderSamplesCounted(sp, cam, W, H, cfg, si, L, pix, nthreads, guide, nullptr);
}

static uint64_t renderSamplesCounted(void* sp, const crh_camera* cam, uint32_t W, uint32_t H,
                                     const crh_integrator_cfg* cfg, uint32_t si, float* L,
                                     float* pix, int nthreads, const Guide* guide,
                                     uint64_t* misUnreachable) {
    const Scene& s = *(Scene*)sp;
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();  // this call only
    uint64_t rays = 0, unreach = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : rays, unreach) num_threads(nt)
    for (int64_t y = 0; y < int64_t(H); ++y) {
        Tally tl;
        for (uint32_t x = 0; x < W; ++x) {
            V2 p;
            Spec r = tracePath(s, *cam, W, H, *cfg, x, uint32_t(y), si, p, tl, guide);
            size_t id = size_t(W) * y + x;
            for (int c = 0; c < 3; ++c) L[3 * id + c] = r.v[c];
            if (pix) {
                pix[2 * id] = p.x;
                pix[2 * id + 1] = p.y;
            }
        }
        rays += tl.rays;
        unreach += tl.misUnreachable;
    }
    if (misUnreachable) *misUnreachable = unreach;
    return rays;
}

// Full render of sample indices [first, first + spp) into (sum, wsum), which
// the caller zeroes.  Deterministic: per-pass radiance is computed in
// parallel, then splatted in row-major pixel order.
static uint64_t renderFilm(void* sp, const crh_camera* cam, uint32_t W, uint32_t H,
                           const crh_integrator_cfg* cfg, uint32_t first, uint32_t spp,
                           float* sum, float* wsum, int nthreads, const Guide* guide) {
    std::vector<float> L(size_t(W) * H * 3), P(size_t(W) * H * 2);
    uint64_t rays = 0;
    for (uint32_t i = 0; i < spp; ++i) {
        rays += renderSamples(sp, cam, W, H, cfg, first + i, L.data(), P.data(), nthreads,
                              guide);
        for (size_t id = 0; id < size_t(W) * H; ++id)
            splat(sum, wsum, W, H, Spec(L[3 * id], L[3 * id + 1], L[3 * id + 2]),
                  V2(P[2 * id], P[2 * id + 1]));
    }
    return rays;
}

uint64_t orc_render(void* sp, const crh_camera* cam, uint32_t W, uint32_t H,
                    const crh_integrator_cfg* cfg, uint32_t first, uint32_t spp,
                    float* sum, float* wsum, int nthreads) {
    return renderFilm(sp, cam, W, H, cfg, first, spp, sum, wsum, nthreads, nullptr);
}

// ---- guided rendering and its pieces (an addition; see Guide) ---------------
// out48: the guide network's 48 outputs, one mixture for every vertex (a
// network whose weights are zero outputs its last bias for any input); NULL
// with bsdf_as_guide != 0 makes the proposal the BSDF's own procedure.
static Guide makeGuide(const float* out48, float alpha, uint32_t mask, uint32_t maxDepth,
                       int bsdfAsGuide) {
    Guide g;
    if (out48) g.mix = mixtureFromOutputs(out48);
    g.alpha = alpha;
    g.mask = mask;
    g.maxDepth = maxDepth;
    g.bsdfAsGuide = bsdfAsGuide != 0 || !out48;
    return g;
}

uint64_t orc_render_guided(void* sp, const crh_camera* cam, uint32_t W, uint32_t H,
                           const crh_integrator_cfg* cfg, const float* out48, float alpha,
                           uint32_t mask, uint32_t maxDepth, int bsdfAsGuide, uint32_t first,
                           uint32_t spp, float* sum, float* wsum, int nthreads) {
    const Guide g = makeGuide(out48, alpha, mask, maxDepth, bsdfAsGuide);
    return renderFilm(sp, cam, W, H, cfg, first, spp, sum, wsum, nthreads, &g);
}

uint64_t orc_render_samples_guided(void* sp, const crh_camera* cam, uint32_t W, uint32_t H,
                                   const crh_integrator_cfg* cfg, const float* out48,
                                   float alpha, uint32_t mask, uint32_t maxDepth,
                                   int bsdfAsGuide, uint32_t si, float* L, float* pix,
                                   int nthreads) {
    const Guide g = makeGuide(out48, alpha, mask, maxDepth, bsdfAsGuide);
    return renderSamples(sp, cam, W, H, cfg, si, L, pix, nthreads, &g);
}

// Mixture of 48 outputs as 40 floats: weights (8), means (8 x 3), kappas (8).
void orc_mixture(const float* out48, float* mix40) {
    Mixture m = mixtureFromOutputs(out48);
    for (int j = 0; j < 8; ++j) {
        mix40[j] = m.w[j];
        for (int c = 0; c < 3; ++c) mix40[8 + 3 * j + c] = m.mu[j][c];
        mix40[32 + j] = m.k[j];
    }
}
float orc_mixture_pdf(const float* out48, const float* d) {
    return mixturePdf(mixtureFromOutputs(out48), V3(d[0], d[1], d[2]));
}
void orc_mixture_sample(const float* out48, const float* u3, float* d) {
    V3 r = mixtureSample(mixtureFromOutputs(out48), u3[0], u3[1], u3[2]);
    for (int c = 0; c < 3; ++c) d[c] = r[c];
}
float orc_actual_pdf_local(const crh_material* m, const float* wo, const float* wi) {
    BSDF b;
    b.init(*m, Frame());
    return actualPdfLocal(b, V3(wo[0], wo[1], wo[2]), V3(wi[0], wi[1], wi[2]));
}
// Batches for the density pins: n BSDF sample directions from u (2 per
// sample), and the actual density at n directions.
void orc_bsdf_sample_dirs(const crh_material* m, const float* wo, const float* u, uint64_t n,
                          float* wi) {
    BSDF b;
    b.init(*m, Frame());
    const V3 o(wo[0], wo[1], wo[2]);
    for (uint64_t i = 0; i < n; ++i) {
        float pdf = 0.f;
        V3 w = b.sampleLocal(o, V2(u[2 * i], u[2 * i + 1]), pdf);
        for (int c = 0; c < 3; ++c) wi[3 * i + c] = w[c];
    }
}
void orc_actual_pdf_batch(const crh_material* m, const float* wo, const float* wi, uint64_t n,
                          float* out) {
    BSDF b;
    b.init(*m, Frame());
    const V3 o(wo[0], wo[1], wo[2]);
    for (uint64_t i = 0; i < n; ++i)
        out[i] = actualPdfLocal(b, o, V3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]));
}
// One guided continuation in the local frame (identity frame) from the
// uniforms u (4 floats, drawn in order); returns wi, f * |cos| and the pdf.
void orc_guided_sample_local(const crh_material* m, const float* out48, float alpha,
                             const float* wo, const float* u, float* wi, float* f, float* pdf) {
    BSDF b;
    b.init(*m, Frame());
    const Guide g = makeGuide(out48, alpha, 3u, 0u, 0);
    FixedUniforms fu{u};
    V3 w;
    float p = 0.f;
    Spec r = guidedSample(b, g, fu, V3(wo[0], wo[1], wo[2]), w, p);
    for (int c = 0; c < 3; ++c) {
        wi[c] = w[c];
        f[c] = r.v[c];
    }
    *pdf = p;
}

int orc_max_threads(void) { return omp_get_max_threads(); }

void orc_cie_cmf(float* xyz) {
    double c[3];
    for (int i = 0; i < 95; ++i) {
        cieXYZ(i, c);
        for (int j = 0; j < 3; ++j) xyz[3 * i + j] = float(c[j]);
    }
}
void orc_spectrum_to_xyz(const float* s, float* xyz) { spectrumToXYZ(s, xyz); }

// One sphere test (SphereSOA::intersect + intersectShape's range test) on a
// one-sphere scene; out = (t, u, v).  Returns 1 on a hit.
int orc_sphere_hit(const float* center, float radius, const crh_ray* r, float* out) {
    Scene s;
    s.sphC.push_back(V3(center[0], center[1], center[2]));
    s.sphR.push_back(radius);
    RayHit rh;
    rh.ray = toRay(*r);
    if (!intersectSphereShape(s, 0, rh) || rh.prim == CRH_INVALID_PRIM) return 0;
    out[0] = rh.hitT;
    out[1] = rh.bary.x;
    out[2] = rh.bary.y;
    return 1;
}

// TextureView::sample of one image at (u, v); out = 4 floats.
void orc_texture_sample(const crh_texture* tx, float u, float v, float* out) {
    Scene::Tex o{tx->width, tx->height,
                 std::vector<float>(4 * size_t(tx->width) * tx->height, 0.f)};
    for (size_t i = 0; i < size_t(tx->width) * tx->height; ++i)
        for (uint32_t c = 0; c < tx->channels; ++c) o.px[4 * i + c] = tx->data[i * tx->channels + c];
    F4 r = textureSample(o, V2(u, v));
    for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}

}  // extern "C"
