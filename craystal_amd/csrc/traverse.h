// Closest-hit and any-hit BVH traversal for gfx950.
//
// Semantics are those of BVHData::intersect (Core/BVH.cu:13-84) with
// intersectShape (Scene/Shape.h:17-37): root-box reject; pop entries whose
// entry t exceeds the closest hit; at an internal node test both child boxes
// and visit the nearer first (ties: left first); push a child only if its
// entry t is below the closest hit; leaf primitives are tested in leaf order
// and a hit replaces the closest one only if strictly nearer.
//
// The reference pushes both children and immediately pops the near one; here
// the near child stays in registers and only the far child is pushed.  The
// popped near child would never fail the `entry.tMin > closestT` test (it was
// pushed with t < closestT and closestT has not changed), so the visit order
// and every comparison are identical.  With that, the stack never holds more
// entries than the tree has levels.  It lives in LDS, entry k of a thread at
// stack[k * stride] with stride = threads per block so a wave's lanes hit
// consecutive 8-byte slots (LdsStack), in private memory (PrivStack), or the
// first kShortStack levels in LDS and the rest in private memory (ShortStack).
#pragma once
#include "device_math.h"
#include "traverse_limits.h"
#include "wave.h"

namespace crh {

struct Hit {
    uint32_t prim;
    float t, u, v;
};

struct WorkCount {
    uint32_t nodes, tris;
};

// Stack entries in LDS: entry k of a thread at slot k * (threads per block)
// + thread of the kernel's dynamic LDS (which starts with the stack), so a
// wave's lanes touch consecutive 8-byte slots.  The slots are addressed
// through an LDS (address space 3) pointer computed from the thread and
// block size at each access: the accesses compile to ds_read / ds_write,
// and no stack object holds a base or stride that would have to live (and
// be reloaded) in private memory next to a private array.
typedef __attribute__((address_space(3))) unsigned long long LdsU64;
extern __shared__ uint2 crhStackLds[];
CRH_D LdsU64* ldsEntry(int k) {
    return (LdsU64*)crhStackLds + (threadIdx.x + uint32_t(k) * blockDim.x);
}
CRH_D void ldsPut(int k, uint2 v) { *ldsEntry(k) = __builtin_bit_cast(unsigned long long, v); }
CRH_D uint2 ldsGet(int k) { return __builtin_bit_cast(uint2, (unsigned long long)*ldsEntry(k)); }

// Traversal stack in LDS.
struct LdsStack {
    CRH_D void put(int k, uint2 v) { ldsPut(k, v); }
    CRH_D uint2 get(int k) const { return ldsGet(k); }
};

// Traversal stack in private (scratch) memory: no LDS, so occupancy is set
// by registers alone; entries are cached in L1/L2.  Holds kPrivStack levels.
constexpr int kPrivStack = kPrivStackLevels;
struct PrivStack {
    uint2 a[kPrivStack];
    CRH_D void put(int k, uint2 v) { a[k] = v; }
    CRH_D uint2 get(int k) const { return a[k]; }
};

// Short stack: levels below kShortStack in LDS (laid out as LdsStack), deeper
// levels in private memory.  Typical
// traversals stay within the LDS part, so the private part (and its scratch
// traffic) is touched only by deep paths; the LDS part is small enough not to
// limit occupancy.
constexpr int kShortStack = kShortStackLds;
struct ShortStack {
    uint2 a[kPrivStack - kShortStack];
    CRH_D void put(int k, uint2 v) {
        if (k < kShortStack) ldsPut(k, v);
        else a[k - kShortStack] = v;
    }
    CRH_D uint2 get(int k) const { return k < kShortStack ? ldsGet(k) : a[k - kShortStack]; }
};

// SphereSOA::intersect (Scene/Sphere.cu:6-40) with intersectShape's range
// test (Scene/Shape.h:17-37): the nearer root if it lies in [tMin, tMax],
// else the farther one.  (u, v) are not computed here; shading derives them
// from the hit point when a texture needs them.
CRH_D bool sphereHit(const float* sph, const Ray& ray, float& t) {
    const float4 q = *reinterpret_cast<const float4*>(sph);
    F3 oc = ray.o - mk3(q.x, q.y, q.z);
    float a = dot(ray.d, ray.d);
    float b = 2.0f * dot(oc, ray.d);
    float c = dot(oc, oc) - q.w * q.w;
    float disc = b * b - 4.0f * a * c;
    if (disc < 0) return false;
    float sq = sqrtf(disc);
    float t0 = (-b - sq) / (2.0f * a);
    float t1 = (-b + sq) / (2.0f * a);
    if (t0 > t1) {
        float x = t0;
        t0 = t1;
        t1 = x;
    }
    if (t0 < ray.tMin || t0 > ray.tMax) {
        if (t1 < ray.tMin || t1 > ray.tMax) return false;
        t0 = t1;
    }
    t = t0;
    return true;
}

// The sphere loop SceneView::intersect runs before the BVH
// (Scene/Scene.cu:5-29): every sphere in index order, a hit replacing the
// current one only if strictly nearer.  ANY: true at the first hit.
template <bool ANY>
CRH_D bool sphereLoop(const SceneView& s, const Ray& ray, Hit& hit) {
    bool isHit = false;
    for (uint32_t i = 0; i < s.nSphere; ++i) {
        float t;
        if (sphereHit(s.spheres + 4 * size_t(i), ray, t)) {
            isHit = true;
            if (ANY) return true;
            if (t < hit.t) {
                hit.t = t;
                hit.u = 0.f;
                hit.v = 0.f;
                hit.prim = i | kSphereBit;
            }
        }
    }
    return isHit;
}

CRH_D void loadNode(const GNode* nodes, uint32_t i, float4& a, float4& b,
                    float4& c, uint4& d) {
    const float4* p = reinterpret_cast<const float4*>(nodes + i);
    a = p[0];
    b = p[1];
    c = p[2];
    d = reinterpret_cast<const uint4*>(p)[3];
}

// Returns true if any primitive was hit inside [tMin, tMax] (the reference's
// return value).  ANY: stop at the first such primitive.
// Both child boxes of a node.  One branch per node: where either box is flat
// (or the ray's range is NaN) both take the reference's slab test boxHit,
// which for a box without flat axes gives boxHitFast's result and t bit for
// bit, so a node's outcome is the same whichever branch it takes.
CRH_D void nodeBoxes(uint32_t fl, const float* lmin, const float* lmax, const float* rmin,
                     const float* rmax, const RayPre& r, float& lt, float& rt, bool& hl,
                     bool& hr) {
    if (fl & 3u) {
        hl = boxHit(lmin, lmax, r, lt);
        hr = boxHit(rmin, rmax, r, rt);
    } else {
        hl = boxHitFast(lmin, lmax, r, lt);
        hr = boxHitFast(rmin, rmax, r, rt);
    }
}

template <bool ANY, bool COUNT, class Stack>
CRH_D bool traverse(const SceneView& s, const Ray& ray, Hit& hit, Stack& stack,
                    WorkCount& wc) {
    hit.prim = 0xFFFFFFFFu;
    hit.t = __builtin_inff();
    hit.u = 0.f;
    hit.v = 0.f;
    bool isHit = false;
    if (s.nSphere) {
        isHit = sphereLoop<ANY>(s, ray, hit);
        if (ANY && isHit) return true;
    }
    if (!s.hasGeom) return isHit;
    RayPre r = prep(ray);
    float dummy;
    if (!boxHit(s.rootMin, s.rootMax, r, dummy)) return isHit;

    float closestT = hit.t;
    uint32_t cur = s.rootRef;
    int sp = 0;
    while (true) {
        // The whole leaf in an inner loop.  (The persistent trace kernel's
        // travStep takes one triangle per step instead, so that lanes in long
        // leaves do not hold their wave; here, where each lane carries a whole
        // path, the inner loop measured faster.)
        if (cur & kLeafBit) {
            uint32_t i = cur & ~kLeafBit;
            while (true) {
                const float4* tp = reinterpret_cast<const float4*>(s.tris + i);
                float4 a = tp[0], b = tp[1], c = tp[2];
                uint32_t pw = __float_as_uint(a.w);
                if (COUNT) wc.tris++;
                float u, v, t;
                if (triHit(mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(c.x, c.y, c.z), r, u, v,
                           t)) {
                    if (!(t < r.tMin || t > r.tMax)) {
                        isHit = true;
                        if (t < hit.t) {
                            hit.t = t;
                            hit.u = u;
                            hit.v = v;
                            hit.prim = pw & ~kLeafBit;
                        }
                        closestT = smin(closestT, hit.t);
                        if (ANY) return true;
                    }
                }
                if (pw & kLeafBit) break;  // the leaf's last triangle
                ++i;
            }
        } else {
            if (COUNT) wc.nodes++;
            float4 a, b, c;
            uint4 d;
            loadNode(s.nodes, cur, a, b, c, d);
            const float lmin[3] = {a.x, a.y, a.z};
            const float lmax[3] = {a.w, b.x, b.y};
            const float rmin[3] = {b.z, b.w, c.x};
            const float rmax[3] = {c.y, c.z, c.w};
            float lt = __builtin_inff(), rt = __builtin_inff();
            bool hl, hr;
            nodeBoxes(slowBoxes(d.z, r), lmin, lmax, rmin, rmax, r, lt, rt, hl, hr);
            if (hl && hr) {
                // Reference pushes the far child first, then the near one
                // (Core/BVH.cu:60-72); near = right iff lt > rt.
                bool rightNear = lt > rt;
                uint32_t nearRef = rightNear ? d.y : d.x;
                uint32_t farRef = rightNear ? d.x : d.y;
                float nearT = rightNear ? rt : lt;
                float farT = rightNear ? lt : rt;
                if (nearT < closestT) {
                    if (farT < closestT) {
                        stack.put(sp, make_uint2(farRef, __float_as_uint(farT)));
                        ++sp;
                    }
                    cur = nearRef;
                    continue;
                }
            } else if (hl && lt < closestT) {
                cur = d.x;
                continue;
            } else if (hr && rt < closestT) {
                cur = d.y;
                continue;
            }
        }
        // Pop the next entry that is not beyond the closest hit.
        bool found = false;
        while (sp > 0) {
            --sp;
            uint2 e = stack.get(sp);
            if (__uint_as_float(e.y) > closestT) continue;
            cur = e.x;
            found = true;
            break;
        }
        if (!found) break;
    }
    return isHit;
}

// ---------------------------------------------------------------------------
// The same traversal as a resumable state machine: travInit + repeated
// travStep perform exactly the operations of traverse(), one internal node or
// one triangle per step, so a persistent kernel can interleave rays per lane
// without changing any result.
// ---------------------------------------------------------------------------
struct TravState {
    RayPre r;
    Hit hit;
    float closestT;
    uint32_t cur;
    int sp;
    bool isHit;
};

// Returns true if the ray is already finished (empty scene / root miss /
// any-hit on a sphere).
template <bool ANY>
CRH_D bool travInit(const SceneView& s, const Ray& ray, TravState& t) {
    t.hit.prim = 0xFFFFFFFFu;
    t.hit.t = __builtin_inff();
    t.hit.u = 0.f;
    t.hit.v = 0.f;
    t.isHit = false;
    if (s.nSphere) {
        t.isHit = sphereLoop<ANY>(s, ray, t.hit);
        if (ANY && t.isHit) return true;
    }
    t.closestT = t.hit.t;
    t.sp = 0;
    t.cur = s.rootRef;
    if (!s.hasGeom) return true;
    t.r = prep(ray);
    float dummy;
    return !boxHit(s.rootMin, s.rootMax, t.r, dummy);
}

// One step of traverse()'s work (one internal node, or up to `tris`
// consecutive triangles of a leaf); returns true when the ray is done.
template <bool ANY, bool COUNT, class Stack>
CRH_D bool travStep(const SceneView& s, TravState& t, Stack& stack, WorkCount& wc,
                    uint32_t tris = 1) {
    const RayPre& r = t.r;
    if (t.cur & kLeafBit) {
        uint32_t i = t.cur & ~kLeafBit;
        bool last = false;
        for (uint32_t k = 0; k < tris; ++k) {
            const float4* tp = reinterpret_cast<const float4*>(s.tris + i);
            float4 a = tp[0], b = tp[1], c = tp[2];
            uint32_t pw = __float_as_uint(a.w);
            if (COUNT) wc.tris++;
            float u, v, tt;
            if (triHit(mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(c.x, c.y, c.z), r, u, v,
                       tt)) {
                if (!(tt < r.tMin || tt > r.tMax)) {
                    t.isHit = true;
                    if (tt < t.hit.t) {
                        t.hit.t = tt;
                        t.hit.u = u;
                        t.hit.v = v;
                        t.hit.prim = pw & ~kLeafBit;
                    }
                    t.closestT = smin(t.closestT, t.hit.t);
                    if (ANY) return true;
                }
            }
            last = (pw & kLeafBit) != 0;  // the leaf's last triangle
            if (last) break;
            ++i;
        }
        if (!last) {
            t.cur = i | kLeafBit;
            return false;
        }
    } else {
        if (COUNT) wc.nodes++;
        float4 a, b, c;
        uint4 d;
        loadNode(s.nodes, t.cur, a, b, c, d);
        const float lmin[3] = {a.x, a.y, a.z};
        const float lmax[3] = {a.w, b.x, b.y};
        const float rmin[3] = {b.z, b.w, c.x};
        const float rmax[3] = {c.y, c.z, c.w};
        float lt = __builtin_inff(), rt = __builtin_inff();
        // Per box here: one branch per node measured 5 % slower in the
        // persistent trace kernel (it is faster in traverse()).
        const uint32_t fl = slowBoxes(d.z, r);
        bool hl = (fl & 1u) ? boxHit(lmin, lmax, r, lt) : boxHitFast(lmin, lmax, r, lt);
        bool hr = (fl & 2u) ? boxHit(rmin, rmax, r, rt) : boxHitFast(rmin, rmax, r, rt);
        if (hl && hr) {
            bool rightNear = lt > rt;
            uint32_t nearRef = rightNear ? d.y : d.x;
            uint32_t farRef = rightNear ? d.x : d.y;
            float nearT = rightNear ? rt : lt;
            float farT = rightNear ? lt : rt;
            if (nearT < t.closestT) {
                if (farT < t.closestT) {
                    stack.put(t.sp, make_uint2(farRef, __float_as_uint(farT)));
                    ++t.sp;
                }
                t.cur = nearRef;
                return false;
            }
        } else if (hl && lt < t.closestT) {
            t.cur = d.x;
            return false;
        } else if (hr && rt < t.closestT) {
            t.cur = d.y;
            return false;
        }
    }
    while (t.sp > 0) {
        --t.sp;
        uint2 e = stack.get(t.sp);
        if (__uint_as_float(e.y) > t.closestT) continue;
        t.cur = e.x;
        return false;
    }
    return true;
}

// Persistent traversal with per-lane ray refill.  Rays [0, n) are handed
// out by a WaveDispenser over *counter (zero at launch), one atomic per
// chunk.  A wave keeps stepping its lanes and refills once `refill` or more
// of them are idle while rays remain; a step tests up to `tris` triangles of
// a leaf.  src.load(i) gives ray i, src.store(i, isHit, hit) its result.
template <bool ANY, bool COUNT, class Stack, class Src>
CRH_D void persistentTrace(const SceneView& s, Src& src, uint32_t* counter, uint32_t n,
                           uint32_t refill, Stack& stack, WorkCount& wc, uint32_t tris = 1) {
    const uint32_t kNone = 0xFFFFFFFFu;
    const uint32_t chunk = dispenserChunk(n);
    uint32_t mine = kNone;
    WaveDispenser disp;
    TravState t;
    while (true) {
        const uint32_t idx = disp.take(counter, n, chunk, mine == kNone);
        if (idx < n) {
            mine = idx;
            if (travInit<ANY>(s, src.load(idx), t)) {
                src.store(idx, t.isHit, t.hit);
                mine = kNone;
            }
        }
        if (__ballot(mine != kNone) == 0) {
            if (disp.exhausted) break;
            continue;  // every lane finished at init: fetch again
        }
        while (true) {
            if (mine != kNone) {
                if (travStep<ANY, COUNT>(s, t, stack, wc, tris)) {
                    src.store(mine, t.isHit, t.hit);
                    mine = kNone;
                }
            }
            uint64_t idle = __ballot(mine == kNone);
            int nIdle = __popcll(idle);
            if (nIdle == 64) break;
            if (!disp.exhausted && nIdle >= int(refill)) break;
        }
    }
}

// SceneView::intersect as the integrator uses it.  A query that reports a
// hit without recording a primitive (only possible when every comparison
// sees a NaN, e.g. a NaN ray direction) is treated as a miss: the reference
// would go on to read an uninitialised primitive index.
template <bool COUNT, class Stack>
CRH_D bool closestHit(const SceneView& s, const Ray& ray, Hit& hit,
                      Stack& stack, WorkCount& wc) {
    bool h = traverse<false, COUNT>(s, ray, hit, stack, wc);
    return h && hit.prim != 0xFFFFFFFFu;
}

// Stack selection: SM = 0 LDS stack, 1 private (scratch) stack, 2 short LDS
// stack with private overflow.
template <int SM>
struct StackOf;
template <>
struct StackOf<0> {
    using T = LdsStack;
};
template <>
struct StackOf<1> {
    using T = PrivStack;
};
template <>
struct StackOf<2> {
    using T = ShortStack;
};

// The stack of the calling thread; `lds` is the kernel's dynamic LDS, whose
// start the LDS stacks occupy (crhStackLds is the same memory).
template <int SM>
CRH_D typename StackOf<SM>::T makeStack(uint2* lds) {
    (void)lds;
    typename StackOf<SM>::T s;  // default-initialised: the private entries are not cleared
    return s;
}

}  // namespace crh
