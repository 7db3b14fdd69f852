// Per-path steps of pathTraceKernel (Core/Integrator.cu:13-227) on a
// thread's own traversal stack, shared by the megakernels and the
// wavefront's finishing kernel: closest / occlusion queries that count the
// rays they trace, sampleLight, evalMIS and one bounce of the path loop.
#pragma once
#include "integrator.h"

namespace crh {

template <bool COUNT, class Stack>
struct PathCtx {
    const SceneView& s;
    Stack& stack;
    uint32_t rays;
    WorkCount wc;
};

template <bool COUNT, class Stack>
CRH_D bool closest(PathCtx<COUNT, Stack>& c, const Ray& r, Hit& h) {
    c.rays++;
    return closestHit<COUNT>(c.s, r, h, c.stack, c.wc);
}
template <bool COUNT, class Stack>
CRH_D bool occluded(PathCtx<COUNT, Stack>& c, const Ray& r) {
    Hit h;
    c.rays++;
    return traverse<true, COUNT>(c.s, r, h, c.stack, c.wc);
}

// sampleLight (Core/Integrator.cu:21-72)
template <bool COUNT, class Stack>
CRH_D LightSample sampleLight(PathCtx<COUNT, Stack>& c, const Isect& it,
                              const BSDF& bsdf, Sampler& smp, const Lanes& w) {
    LightGeom g = lightGeom(c.s, it, smp);
    if (!g.valid) return noLight();
    if (occluded(c, g.shadow)) return noLight();
    return lightFinish(c.s, g, bsdf, it, w);
}

// evalMIS (Core/Integrator.cu:74-140)
template <bool COUNT, class Stack>
CRH_D Spec evalMIS(PathCtx<COUNT, Stack>& c, const Isect& it, const BSDF& bsdf,
                   Sampler& smp, const Lanes& w) {
    LightSample ls = sampleLight(c, it, bsdf, smp, w);
    if (ls.pdf <= 0.0f) return sp(0.f);
    Spec value = misLightPart(ls, bsdf, it);
    Ray r;
    Spec bw;
    float bp;
    if (misBsdfRay(bsdf, smp, it, r, bw, bp) && mayHitEmitter(c.s, r, ls.id)) {
        Hit h;
        bool isHit = closest(c, r, h);
        misBsdfPart(c.s, it.pos, r, isHit, h, ls.id, bw, bp, value, w);
    }
    return value / ls.selPdf;
}

// One iteration of pathTraceKernel's bounce loop (:160-223) at `depth`:
// extension hit, emitter or environment term, light sample (with MIS),
// continuation sample, the unconditional kill and roulette.  Returns false
// when the path ends; the caller stops at maxDepth.  SPEC: spectral lanes.
template <bool MIS, bool SPEC, bool COUNT, class Stack>
CRH_D bool pathBounce(PathCtx<COUNT, Stack>& c, const IntegratorView& cfg, uint32_t depth,
                      Ray& ray, Spec& rad, Spec& beta, Sampler& smp, const Lanes& w) {
    const SceneView& s = c.s;
    bool terminate = false;
    Hit h;
    if (closest(c, ray, h)) {
        Isect it = makeIsect(s, ray, h);
        const uint32_t mid = matIdOf(s, h.prim);
        const GMaterial& m = s.mats[mid];
        if (m.flags & 1u) {
            rad = rad + emissionOf(m, mid, w) * beta;
            return false;
        }
        BSDF bsdf = makeBSDF(s, m, mid, it, h, w);
        if (MIS) {
            rad = rad + evalMIS(c, it, bsdf, smp, w) * beta;
        } else {
            LightSample ls = sampleLight(c, it, bsdf, smp, w);
            if (ls.pdf > 0.0f) rad = rad + ls.weight * beta / ls.pdf / ls.selPdf;
        }
        F3 wi;
        float pdf = 0.0f;
        Spec f = bsdfSampleEval(bsdf, smp, it.view, wi, pdf);
        if (pdf <= 0.0f) return false;
        beta = beta * (f / pdf);
        ray.o = offsetRayOrigin(it.pos, oriented(it));
        ray.d = wi;
    } else {
        rad = rad + (SPEC ? laneSpec(s.spec + kSpecEnv, w) : sp3(s.env)) * beta;
        terminate = true;
    }
    if (smp.next1D() < cfg.rrThreshold) terminate = true;
    if (terminate) return false;
    return roulette(cfg, depth, smp, beta);
}

}  // namespace crh
