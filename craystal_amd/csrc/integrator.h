// Per-path integrator steps shared by the megakernel and the wavefront
// integrator, split at the points where a ray query happens so that both
// schedules execute the identical arithmetic in the identical order:
//
//   sampleLight (Core/Integrator.cu:21-72)  = lightGeom -> [shadow query] -> lightFinish
//   evalMIS     (Core/Integrator.cu:74-140) = misLightPart -> misBsdfRay ->
//                                              [closest query] -> misBsdfPart
//                (the query skipped where mayHitEmitter shows it cannot
//                 reach the sampled emitter: misBsdfPart would add nothing)
//
// Random numbers are drawn only inside lightGeom (3: select + barycentric),
// misBsdfRay (2) and the continuation sample, in the reference's order.
#pragma once
#include "shading.h"

namespace crh {

struct LightGeom {
    Ray shadow;
    F3 ld, ln;
    float d2, area;
    uint32_t eid;
    bool valid;
};

struct LightSample {
    Spec weight;
    F3 dir;
    float pdf;
    uint32_t id;
    float selPdf;
};

CRH_D LightSample noLight() {
    LightSample ls;
    ls.weight = sp(0.f);
    ls.dir = mk3(0, 0, 0);
    ls.pdf = 0.f;
    ls.id = 0xFFFFFFFFu;
    ls.selPdf = 0.f;
    return ls;
}

// Light selection, point on the emitter and the validity tests before the
// shadow query (Core/Integrator.cu:25-53).
CRH_D LightGeom lightGeom(const SceneView& s, const Isect& it, Sampler& smp) {
    LightGeom g;
    g.valid = false;
    const int cnt = int(s.nEmissive);
    if (cnt == 0) return g;
    int sel = int(float(cnt) * smp.next1D());
    sel = sel < cnt - 1 ? sel : cnt - 1;
    g.eid = s.emissive[sel];
    TriData tri = getTri(s, g.eid);
    F2 bc = sampleBary(smp.next2D());
    F3 w = mk3(1.f - bc.x - bc.y, bc.x, bc.y);
    F3 lp = interp3(tri.p, w);
    g.ln = interp3(tri.n, w);
    F3 target = offsetRayOrigin(lp, faceNormal(tri));
    F3 toL = target - it.pos;
    g.d2 = dot(toL, toL);
    float dist = sqrtf(g.d2);
    g.ld = toL / dist;
    F3 on = oriented(it);
    if (dot(on, g.ld) < 0.0f || dot(g.ln, g.ld) > 0.0f || dist < kEps) return g;
    g.shadow.o = offsetRayOrigin(it.pos, on);
    g.shadow.tMin = 0.f;
    g.shadow.d = g.ld;
    g.shadow.tMax = length(g.shadow.o - target) - kEps;
    g.area = triArea(tri);
    g.valid = true;
    return g;
}

// The unoccluded light sample (Core/Integrator.cu:57-70).
CRH_D LightSample lightFinish(const SceneView& s, const LightGeom& g,
                              const BSDF& bsdf, const Isect& it,
                              const Lanes& w = rgbLanes()) {
    LightSample ls;
    const uint32_t mid = s.primMat[g.eid];
    const GMaterial& m = s.mats[mid];
    ls.weight = bsdfEval(bsdf, it.view, g.ld) * emissionOf(m, mid, w);
    ls.dir = g.ld;
    ls.pdf = g.d2 / (g.area * absDot(g.ld, g.ln));
    ls.id = g.eid;
    ls.selPdf = float(1.0 / double(s.nEmissive));
    return ls;
}

// Light half of evalMIS (:79-91); returns the running `value`.
CRH_D Spec misLightPart(const LightSample& ls, const BSDF& bsdf,
                        const Isect& it) {
    Spec value = sp(0.f);
    float lp = ls.pdf;
    float bp = bsdfPdf(bsdf, it.view, ls.dir);
    if (bp > 0.0f) value = value + powerHeur(lp, bp) * ls.weight / lp;
    return value;
}

// BSDF sample of evalMIS (:93-99).  Returns true if the MIS ray is traced.
CRH_D bool misBsdfRay(const BSDF& bsdf, Sampler& smp, const Isect& it, Ray& r,
                      Spec& bw, float& bp) {
    F3 dir;
    bp = 0.0f;
    bw = bsdfSampleEval(bsdf, smp, it.view, dir, bp);
    if (!(smaxv(bw) > 0.0f && bp > 0.0f)) return false;
    r.o = offsetRayOrigin(it.pos, oriented(it));
    r.tMin = 0.f;
    r.d = dir;
    r.tMax = __builtin_inff();
    return true;
}

// Whether the MIS BSDF ray r can have the sampled emitter triangle eid as its
// closest hit.  The traversal records a triangle only where the
// Moller-Trumbore test on it (triHit, on the same vertex differences the
// BVH's triangle records hold) gives a t inside the ray's range and below
// the closest hit so far (at most +inf), so where that test fails on
// triangle eid the closest hit is another primitive or none, the emitter
// half of evalMIS adds nothing, and the ray need not be traced.  With
// analytic spheres in the scene a sphere whose index equals eid matches too
// (hitsLight), so there every ray is traced.
CRH_D bool mayHitEmitter(const SceneView& s, const Ray& r, uint32_t eid) {
    if (s.nSphere) return true;
    const TriData tri = getTri(s, eid);
    RayPre p;
    p.o = r.o;
    p.d = r.d;
    p.tMin = r.tMin;
    p.tMax = r.tMax;
    float u, v, t;
    return triHit(tri.p[0], tri.p[1] - tri.p[0], tri.p[2] - tri.p[0], p, u, v, t) &&
           !(t < r.tMin || t > r.tMax) && t < __builtin_inff();
}

// Emitter half of evalMIS (:100-135) given the MIS ray's closest-hit query
// result (`isHit` = the reference's intersect() return).  Adds into value.
CRH_D void misBsdfPart(const SceneView& s, F3 pos, const Ray& r, bool isHit,
                       const Hit& h, uint32_t eid, Spec bw, float bp,
                       Spec& value, const Lanes& w = rgbLanes()) {
    if (!(isHit && hitsLight(s, h.prim, eid))) return;
    Isect lit = makeIsect(s, r, h);
    const uint32_t mid = matIdOf(s, h.prim);
    const GMaterial& m = s.mats[mid];
    if (!(m.flags & 1u)) return;
    F3 toL = lit.pos - pos;
    float d2 = dot(toL, toL);
    F3 lv = toL / sqrtf(d2);
    float cl = absDot(lv, lit.faceN);
    if (cl != 0.0f) {
        float area = triArea(getTri(s, eid));  // the matched index's triangle
        float lpdf = d2 / (area * cl);
        value = value + powerHeur(bp, lpdf) * emissionOf(m, mid, w) * bw / bp;
    }
}

// CameraProxy::generateRay (Core/CameraProxy.h:25-40).
// Pixel `inTile` (0..255) of a 16x16 tile; each 64-pixel quarter is an 8x8
// block (one wave of a 256-thread block).
CRH_D void tileCoordOf(uint32_t tile, uint32_t inTile, uint32_t tilesX, uint32_t& x,
                       uint32_t& y) {
    uint32_t tx = tile % tilesX, ty = tile / tilesX;
    uint32_t wave = inTile >> 6, lane = inTile & 63u;
    x = tx * 16u + (wave & 1u) * 8u + (lane & 7u);
    y = ty * 16u + (wave >> 1) * 8u + (lane >> 3);
}

CRH_D Ray cameraRay(const CameraView& cam, F2 pix) {
    float u = (2.0f * (pix.x / float(cam.width)) - 1.0f);
    float v = (1.0f - 2.0f * (pix.y / float(cam.height)));
    float hw = cam.aspect * cam.hh;
    F3 dir = normalize(ld3(cam.u) * (u * hw) + ld3(cam.v) * (v * cam.hh) +
                       ld3(cam.w));
    return Ray{ld3(cam.pos), 0.f, dir, __builtin_inff()};
}

// Russian roulette after the unconditional kill (:209-223).  Returns false
// if the path stops.
CRH_D bool roulette(const IntegratorView& cfg, uint32_t depth, Sampler& smp,
                    Spec& beta) {
    if (depth >= cfg.rrStartDepth) {
        float q = smax(cfg.rrThreshold, 1.0f - smaxv(beta));
        q = q < 0.0f ? 0.0f : (1.0f < q ? 1.0f : q);
        if (smp.next1D() < q) return false;
        beta = beta / (1.0f - q);
    }
    return true;
}

}  // namespace crh
