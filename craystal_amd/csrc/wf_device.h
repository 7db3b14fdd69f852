// Device helpers of the wavefront kernels (wavefront.hip, wf_finish.hip):
// path flags, the structure-of-arrays loads and stores of rays, hits and
// spectra, and the per-path sampler.
#pragma once
#include <hip/hip_runtime.h>

#include "batch_device.h"
#include "integrator.h"
#include "path.h"
#include "wavefront.h"

namespace crh {
namespace wf {

enum : uint32_t {
    F_MIS = 1u,      // MIS ray pending (traced by the next extend)
    F_LIGHT = 2u,    // light geometry valid, shadow ray traced
    F_MISHIT = 4u,   // MIS ray query returned a hit (intersect() == true)
    F_EXTHIT = 8u,   // extension query returned a hit
    F_ALIVE = 16u,   // the path continues (it is in the next extension queue)
};

CRH_D uint32_t laneId() { return __lane_id(); }

CRH_D Ray loadRay(const float4* o, const float4* d, uint32_t p) {
    float4 a = o[p], b = d[p];
    return Ray{mk3(a.x, a.y, a.z), a.w, mk3(b.x, b.y, b.z), b.w};
}
CRH_D void storeRay(float4* o, float4* d, uint32_t p, const Ray& r) {
    o[p] = make_float4(r.o.x, r.o.y, r.o.z, r.tMin);
    d[p] = make_float4(r.d.x, r.d.y, r.d.z, r.tMax);
}
CRH_D Hit loadHit(const uint4* h, uint32_t p) {
    uint4 v = h[p];
    return Hit{v.x, __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w)};
}
CRH_D void storeHit(uint4* h, uint32_t p, const Hit& x) {
    h[p] = make_uint4(x.prim, __float_as_uint(x.t), __float_as_uint(x.u), __float_as_uint(x.v));
}
CRH_D Spec ldS(const float4* a, uint32_t p) {
    float4 v = a[p];
    return Spec{v.x, v.y, v.z};
}
CRH_D void stS(float4* a, uint32_t p, Spec s, float w = 0.f) {
    a[p] = make_float4(s.r, s.g, s.b, w);
}
// Where a path's scene comes from: the one scene of a render (OneScene, the
// view passed by value) or the batch table (SceneTable), in which case the
// path carries its scene's index in bits kSceneShift.. of its flags (set by
// the generate kernel) and pixel p of a pass is pixel p - pixPrefix[j] of
// scene j.
constexpr uint32_t kSceneShift = 8, kSceneMask = 0x1FFFu;  // up to 8192 scenes
CRH_D uint32_t sceneOf(uint32_t fl) { return (fl >> kSceneShift) & kSceneMask; }
struct OneScene {
    SceneView s;
    static constexpr bool kBatch = false;
    CRH_D const SceneView& view(uint32_t) const { return s; }
};
struct SceneTable {
    const BatchScene* tab;
    const uint32_t* pixPrefix;  // first pixel of each active scene, then the total
    uint32_t nActive;
    static constexpr bool kBatch = true;
    CRH_D SceneView view(uint32_t j) const { return tab[j].view; }
};
// The scene index of path p (0 for one scene).
template <class Scenes>
CRH_D uint32_t sceneIdx(const WfState& w, uint32_t p) {
    if constexpr (Scenes::kBatch) return sceneOf(w.flags[p]);
    else return 0u;
}

CRH_D Sampler loadSampler(const WfState& w, uint32_t p) {
    uint32_t pix = p % w.npix;
    uint32_t x = pix % w.width, y = pix / w.width;
    Sampler s(0u, 0u, 0u);
    s.state = w.rng[p];
    s.inc = x * 1234567891u + y * 987654321u + 1u;
    return s;
}

// The sampler of path p of a batched pass: the pixel coordinates are its
// scene's (flags fl carry the scene index).
template <class Scenes>
CRH_D Sampler loadSampler(const Scenes& S, const WfState& w, uint32_t p, uint32_t fl) {
    if constexpr (!Scenes::kBatch) {
        return loadSampler(w, p);
    } else {
        const uint32_t j = sceneOf(fl);
        const uint32_t pix = p % w.npix - S.pixPrefix[j], width = S.tab[j].cam.width;
        const uint32_t x = pix % width, y = pix / width;
        Sampler s(0u, 0u, 0u);
        s.state = w.rng[p];
        s.inc = x * 1234567891u + y * 987654321u + 1u;
        return s;
    }
}

#define PERSISTENT_LOOP(i, count)                                            \
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (count);    \
         i += gridDim.x * blockDim.x)


// Spectral transport: the hero bin of a path rides in the top bits of its
// flags (set by the generate kernel, kept by every later flag update).
constexpr uint32_t kHeroShift = 24;
CRH_D Lanes pathLanes(const SceneView& s, uint32_t fl) {
    const uint32_t h = fl >> kHeroShift;
    return Lanes{s.spec, h, (h + 32u) % kSpecBins, (h + 64u) % kSpecBins};
}

}  // namespace wf
}  // namespace crh
