// Wavefront integrator for gfx950.
//
// The megakernel keeps a whole path (traversal stack, BSDF, light sample) in
// one thread; here the same per-path steps (integrator.h) run as a sequence
// of small kernels over queues of path ids, one bounce at a time:
//
//   generate   camera rays for every pixel of the pass        -> extend queue
//   extend     closest hit for extension rays + MIS rays      (traversal only)
//   misResolve finish the MIS estimate of the previous bounce  (MIS queue)
//   shadeHit   emitter / environment, light selection          -> shadow queue
//   shadow     any-hit for the shadow rays                     (traversal only)
//   shadeCont  light weight, MIS BSDF ray, continuation sample -> next queues
//
// Because the reference draws the MIS BSDF sample only when the shadow ray
// is unoccluded (Core/Integrator.cu:80-99), the continuation sample waits for
// the shadow result; the MIS ray of bounce k is traced together with the
// extension ray of bounce k+1 and resolved before bounce k+1 is shaded,
// which is the order in which the megakernel adds those terms.  Every path
// therefore consumes the same random numbers and performs the same float
// operations as in the megakernel, and the film is bit-identical.
//
// Queues are compacted per wave: a ballot of the lanes that push, one atomic
// per wave on the queue counter (readfirstlane broadcast), mbcnt for each
// lane's slot.  All kernels are persistent: a fixed grid loops over the
// queue length read from device memory, so no host synchronisation is
// needed between bounces.
#include <hip/hip_runtime.h>

#include "guide_device.h"
#include "guiding_kernels.h"
#include "kernels.h"
#include "wave.h"
#include "wf_device.h"

namespace crh {

using namespace wf;

// Counter layout (WfCounters): ext[2], mis[2], shadow, cont.
__global__ __launch_bounds__(64) void kWfReset(uint32_t* c, uint32_t mask) {
    uint32_t i = threadIdx.x;
    if (i < kWfCounters && (mask >> i) & 1u) c[i] = 0u;
}

// SPEC: the path draws its hero bin right after the pixel jitter (as the
// megakernel's tracePixel) and keeps it in its flags.  A batched pass's
// pixel p % npix is pixel p % npix - pixPrefix[j] of active scene j, seen
// through that scene's camera; the path keeps j in its flags.
template <bool SPEC, class Scenes>
__global__ __launch_bounds__(256) void kWfGenerate(WfState w, CameraView cam, uint32_t si,
                                                   Scenes S) {
    uint32_t n = w.active * w.npix;
    PERSISTENT_LOOP(p, n) {
        uint32_t px = p % w.npix, width = w.width, sc = 0u;
        const CameraView* cv = &cam;
        if constexpr (Scenes::kBatch) {
            sc = findScene(S.pixPrefix, S.nActive, px);
            px -= S.pixPrefix[sc];
            cv = &S.tab[sc].cam;
            width = cv->width;
        }
        uint32_t x = px % width, y = px / width;
        Sampler smp(x, y, si + p / w.npix);
        F2 j = smp.next2D();
        F2 pix{float(x) + j.x, float(y) + j.y};
        const uint32_t hero = SPEC ? heroLanes(S.view(sc).spec, smp.next1D()).b0 : 0u;
        Ray r = cameraRay(*cv, pix);
        storeRay(w.rayO, w.rayD, p, r);
        w.rng[p] = smp.state;
        stS(w.beta, p, sp(1.f));
        stS(w.rad, p, sp(0.f));
        w.flags[p] = F_ALIVE | (hero << kHeroShift) | (sc << kSceneShift);
        if (w.recId) w.recId[p] = kNoRecord;
        w.P[p] = make_float2(pix.x, pix.y);
        w.qExt[0][p] = p;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) w.counters[kCExt0] = n;
}

// Closest hit for the extension queue followed by the MIS queue.
template <bool COUNT, int SM, class Scenes>
__global__ __launch_bounds__(256) void kWfExtend(Scenes S, WfState w,
                                                 uint32_t cur) {
    extern __shared__ uint2 lds[];
    auto stk = makeStack<SM>(lds);
    const uint32_t nExt = w.counters[kCExt0 + cur];
    const uint32_t nMis = w.counters[kCMis0 + cur];
    if (blockIdx.x == 0 && threadIdx.x == 0 && nExt + nMis)
        atomicAdd(w.work, (unsigned long long)(nExt + nMis));
    WorkCount wc{0, 0};
    PERSISTENT_LOOP(i, nExt + nMis) {
        Hit h;
        if (i < nExt) {
            uint32_t p = w.qExt[cur][i];
            Ray r = loadRay(w.rayO, w.rayD, p);
            bool hit = closestHit<COUNT>(S.view(sceneIdx<Scenes>(w, p)), r, h, stk, wc);
            storeHit(w.hit, p, h);
            if (hit) atomicOr(&w.flags[p], F_EXTHIT);
        } else {
            uint32_t p = w.qMis[cur][i - nExt];
            Ray r = loadRay(w.misO, w.misD, p);
            bool hit = closestHit<COUNT>(S.view(sceneIdx<Scenes>(w, p)), r, h, stk, wc);
            storeHit(w.misHit, p, h);
            if (hit) atomicOr(&w.flags[p], F_MISHIT);
        }
    }
    if (COUNT) {
        unsigned long long n64 = wc.nodes, t64 = wc.tris;
        for (int off = 32; off > 0; off >>= 1) {
            n64 += __shfl_down(n64, off, 64);
            t64 += __shfl_down(t64, off, 64);
        }
        if (laneId() == 0) {
            atomicAdd(w.work + 1, n64);
            atomicAdd(w.work + 2, t64);
        }
    }
}

// Persistent variants: lanes take the next queue entry when their ray ends.
struct ExtendSrc {
    WfState w;
    uint32_t cur, nExt;
    CRH_D Ray load(uint32_t i) const {
        if (i < nExt) return loadRay(w.rayO, w.rayD, w.qExt[cur][i]);
        return loadRay(w.misO, w.misD, w.qMis[cur][i - nExt]);
    }
    CRH_D void store(uint32_t i, bool isHit, const Hit& h) const {
        bool hit = isHit && h.prim != 0xFFFFFFFFu;  // closestHit()
        if (i < nExt) {
            uint32_t p = w.qExt[cur][i];
            storeHit(w.hit, p, h);
            if (hit) atomicOr(&w.flags[p], F_EXTHIT);
        } else {
            uint32_t p = w.qMis[cur][i - nExt];
            storeHit(w.misHit, p, h);
            if (hit) atomicOr(&w.flags[p], F_MISHIT);
        }
    }
};

template <int SM>
__global__ __launch_bounds__(256) void kWfExtendPT(SceneView s, WfState w, uint32_t cur,
                                                   uint32_t refill) {
    extern __shared__ uint2 lds[];
    auto stk = makeStack<SM>(lds);
    const uint32_t nExt = w.counters[kCExt0 + cur];
    const uint32_t nMis = w.counters[kCMis0 + cur];
    if (blockIdx.x == 0 && threadIdx.x == 0 && nExt + nMis)
        atomicAdd(w.work, (unsigned long long)(nExt + nMis));
    ExtendSrc src{w, cur, nExt};
    WorkCount wc{0, 0};
    persistentTrace<false, false>(s, src, &w.counters[kCPtExt], nExt + nMis, refill, stk, wc);
}

struct ShadowSrc {
    WfState w;
    CRH_D Ray load(uint32_t i) const { return loadRay(w.shO, w.shD, w.qShadow[i]); }
    CRH_D void store(uint32_t i, bool isHit, const Hit&) const {
        if (isHit) w.flags[w.qShadow[i]] &= ~F_LIGHT;
    }
};

template <int SM>
__global__ __launch_bounds__(256) void kWfShadowPT(SceneView s, WfState w, uint32_t refill) {
    extern __shared__ uint2 lds[];
    auto stk = makeStack<SM>(lds);
    const uint32_t n = w.counters[kCShadow];
    if (blockIdx.x == 0 && threadIdx.x == 0 && n) atomicAdd(w.work, (unsigned long long)n);
    ShadowSrc src{w};
    WorkCount wc{0, 0};
    persistentTrace<true, false>(s, src, &w.counters[kCPtShadow], n, refill, stk, wc);
}

// Resolve the MIS estimate of the previous bounce:
// radiance += (value / selPdf) * beta_k  (Core/Integrator.cu:139, :186).
template <bool SPEC, class Scenes>
__global__ __launch_bounds__(256) void kWfMisResolve(Scenes S, WfState w,
                                                     uint32_t cur) {
    const uint32_t n = w.counters[kCMis0 + cur];
    PERSISTENT_LOOP(i, n) {
        uint32_t p = w.qMis[cur][i];
        uint32_t fl = w.flags[p];
        const auto& s = S.view(sceneOf(fl));
        Ray r = loadRay(w.misO, w.misD, p);
        Hit h = loadHit(w.misHit, p);
        float4 vs = w.misVal[p];
        Spec value{vs.x, vs.y, vs.z};
        float4 bw4 = w.misBw[p];
        float4 pe = w.misPos[p];
        misBsdfPart(s, mk3(pe.x, pe.y, pe.z), r, (fl & F_MISHIT) != 0u, h,
                    __float_as_uint(pe.w), Spec{bw4.x, bw4.y, bw4.z}, bw4.w, value,
                    SPEC ? pathLanes(s, fl) : rgbLanes());
        Spec rad = ldS(w.rad, p);
        rad = rad + (value / vs.w) * ldS(w.misBeta, p);
        stS(w.rad, p, rad);
        w.flags[p] = fl & ~(F_MIS | F_MISHIT);
    }
}

// Process the extension hit: environment / emitter terms, or light
// selection producing a shadow ray (Core/Integrator.cu:165-190, :25-53).
template <bool GUIDED, bool SPEC, class Scenes>
__global__ __launch_bounds__(256) void kWfShadeHit(Scenes S, WfState w,
                                                   uint32_t cur) {
    const uint32_t n = w.counters[kCExt0 + cur];
    PERSISTENT_LOOP(i, n) {
        uint32_t p = w.qExt[cur][i];
        uint32_t fl = w.flags[p];
        const auto& s = S.view(sceneOf(fl));
        Spec beta = ldS(w.beta, p);
        if (GUIDED) {
            // Radiance before anything found along the guided direction of
            // the previous vertex: the start of that direction's estimate.
            uint32_t rid = w.recId[p];
            if (rid != kNoRecord) {
                float4 r = w.rad[p];
                w.rec[size_t(rid) * kRecF4 + 6] = make_float4(r.x, r.y, r.z, 1.f);
                w.recId[p] = kNoRecord;
            }
        }
        bool shade = false, shadow = false;
        const Lanes lw = SPEC ? pathLanes(s, fl) : rgbLanes();
        if (!(fl & F_EXTHIT)) {
            const Spec env = SPEC ? laneSpec(s.spec + kSpecEnv, lw) : sp3(s.env);
            stS(w.rad, p, ldS(w.rad, p) + env * beta);
        } else {
            Ray r = loadRay(w.rayO, w.rayD, p);
            Hit h = loadHit(w.hit, p);
            const uint32_t mid = matIdOf(s, h.prim);
            const GMaterial& m = s.mats[mid];
            if (m.flags & 1u) {
                stS(w.rad, p, ldS(w.rad, p) + emissionOf(m, mid, lw) * beta);
            } else {
                shade = true;
                Isect it = makeIsect(s, r, h);
                if (GUIDED) {
                    w.gPos[p] = make_float4(it.pos.x, it.pos.y, it.pos.z, 0.f);
                    w.gNrm[p] = make_float4(it.frame.N.x, it.frame.N.y, it.frame.N.z, 0.f);
                    w.gWo[p] = make_float4(it.view.x, it.view.y, it.view.z, 0.f);
                }
                Sampler smp = loadSampler(S, w, p, fl);
                LightGeom g = lightGeom(s, it, smp);
                w.rng[p] = smp.state;
                if (g.valid) {
                    shadow = true;
                    storeRay(w.shO, w.shD, p, g.shadow);
                    w.lgt[p] = make_float4(g.ld.x, g.ld.y, g.ld.z, g.d2);
                    w.lgn[p] = make_float4(g.ln.x, g.ln.y, g.ln.z, g.area);
                    w.lgId[p] = g.eid;
                    fl |= F_LIGHT;
                } else {
                    fl &= ~F_LIGHT;
                }
            }
        }
        w.flags[p] = fl & ~F_EXTHIT;
        uint32_t slot = wavePush(&w.counters[kCCont], shade);
        if (shade) w.qCont[slot] = p;
        slot = wavePush(&w.counters[kCShadow], shadow);
        if (shadow) w.qShadow[slot] = p;
    }
}

template <int SM, class Scenes>
__global__ __launch_bounds__(256) void kWfShadow(Scenes S, WfState w) {
    extern __shared__ uint2 lds[];
    auto stk = makeStack<SM>(lds);
    const uint32_t n = w.counters[kCShadow];
    if (blockIdx.x == 0 && threadIdx.x == 0 && n) atomicAdd(w.work, (unsigned long long)n);
    WorkCount wc{0, 0};
    PERSISTENT_LOOP(i, n) {
        uint32_t p = w.qShadow[i];
        Ray r = loadRay(w.shO, w.shD, p);
        Hit h;
        bool occ = traverse<true, false>(S.view(sceneIdx<Scenes>(w, p)), r, h, stk, wc);
        if (occ) w.flags[p] &= ~F_LIGHT;  // only this thread touches p here
    }
}

// Light weight, MIS BSDF ray and the continuation sample
// (Core/Integrator.cu:57-99, :186-223).
template <bool MIS, bool GUIDED, bool SPEC, class Scenes>
__global__ __launch_bounds__(256) void kWfShadeCont(Scenes S, WfState w,
                                                    IntegratorView cfg,
                                                    uint32_t cur, uint32_t depth,
                                                    float alpha, uint32_t gmask) {
    const uint32_t n = w.counters[kCCont];
    const uint32_t nxt = cur ^ 1u;
    PERSISTENT_LOOP(i, n) {
        uint32_t p = w.qCont[i];
        uint32_t fl = w.flags[p];
        const auto& s = S.view(sceneOf(fl));
        Ray r = loadRay(w.rayO, w.rayD, p);
        Hit h = loadHit(w.hit, p);
        Isect it = makeIsect(s, r, h);
        const uint32_t mid = matIdOf(s, h.prim);
        const GMaterial& m = s.mats[mid];
        const Lanes lw = SPEC ? pathLanes(s, fl) : rgbLanes();
        BSDF bsdf = makeBSDF(s, m, mid, it, h, lw);
        Sampler smp = loadSampler(S, w, p, fl);
        Spec beta = ldS(w.beta, p);
        Spec rad = ldS(w.rad, p);
        LightSample ls = noLight();
        if (fl & F_LIGHT) {
            float4 a = w.lgt[p], b = w.lgn[p];
            LightGeom g;
            g.ld = mk3(a.x, a.y, a.z);
            g.d2 = a.w;
            g.ln = mk3(b.x, b.y, b.z);
            g.area = b.w;
            g.eid = w.lgId[p];
            g.valid = true;
            ls = lightFinish(s, g, bsdf, it, lw);
        }
        bool misRay = false;
        if (MIS) {
            if (ls.pdf <= 0.0f) {
                rad = rad + sp(0.f) * beta;
            } else {
                Spec value = misLightPart(ls, bsdf, it);
                Ray mr;
                Spec bw;
                float bp;
                if (misBsdfRay(bsdf, smp, it, mr, bw, bp) && mayHitEmitter(s, mr, ls.id)) {
                    misRay = true;
                    storeRay(w.misO, w.misD, p, mr);
                    stS(w.misVal, p, value, ls.selPdf);
                    stS(w.misBw, p, bw, bp);
                    w.misPos[p] = make_float4(it.pos.x, it.pos.y, it.pos.z,
                                              __uint_as_float(ls.id));
                    stS(w.misBeta, p, beta);
                } else {
                    rad = rad + (value / ls.selPdf) * beta;
                }
            }
        } else if (ls.pdf > 0.0f) {
            rad = rad + ls.weight * beta / ls.pdf / ls.selPdf;
        }
        // Continuation (Core/Integrator.cu:195-223).
        bool alive = true;
        F3 wi;
        float pdf = 0.0f;
        Spec f;
        // gmask: bit t set = guide BSDF type t (0 Lambertian, 1 Principled);
        // Phong is never guided (its sampled density is not the mixture's).
        bool guided = GUIDED && bsdf.type != 2u && ((gmask >> bsdf.type) & 1u);
        float qmix = 0.f;
        if (!guided) {
            f = bsdfSampleEval(bsdf, smp, it.view, wi, pdf);
        } else {
            Mixture mix = loadMixture(w.gMix + size_t(p) * 40);
            f = guidedSampleEval(bsdf, mix, alpha, smp, it.view, wi, pdf, qmix);
        }
        bool record = false;
        Spec betaW = beta;
        if (pdf <= 0.0f) {
            alive = false;
        } else {
            beta = beta * (f / pdf);
            betaW = beta;
            record = guided;
            Ray nr{offsetRayOrigin(it.pos, oriented(it)), r.tMin, wi, r.tMax};
            storeRay(w.rayO, w.rayD, p, nr);
            if (smp.next1D() < cfg.rrThreshold) alive = false;
            else if (!roulette(cfg, depth, smp, beta)) alive = false;
        }
        if (depth + 1 >= cfg.maxDepth) alive = false;
        if (GUIDED) {
            uint32_t slot = wavePush(&w.counters[kCRec], record);
            if (record && slot < w.recCap) {
                float4* R = w.rec + size_t(slot) * kRecF4;
                R[0] = make_float4(it.pos.x, it.pos.y, it.pos.z, qmix);
                R[1] = make_float4(it.frame.N.x, it.frame.N.y, it.frame.N.z, 0.f);
                R[2] = make_float4(it.view.x, it.view.y, it.view.z, 0.f);
                R[3] = make_float4(wi.x, wi.y, wi.z, __uint_as_float(p));
                R[4] = make_float4(f.r, f.g, f.b, 0.f);
                R[5] = make_float4(betaW.r, betaW.g, betaW.b, 0.f);  // before roulette
                R[6] = make_float4(0.f, 0.f, 0.f, 0.f);
                w.recId[p] = slot;
            } else {
                w.recId[p] = kNoRecord;
            }
        }
        w.rng[p] = smp.state;
        stS(w.beta, p, beta);
        stS(w.rad, p, rad);
        w.flags[p] = (fl & ~(F_LIGHT | F_ALIVE)) | (misRay ? F_MIS : 0u) | (alive ? F_ALIVE : 0u);
        uint32_t slot = wavePush(&w.counters[kCExt0 + nxt], alive);
        if (alive) w.qExt[nxt][slot] = p;
        slot = wavePush(&w.counters[kCMis0 + nxt], misRay);
        if (misRay) w.qMis[nxt][slot] = p;
    }
}

// Spectral transport: every path's three-lane radiance to CIE XYZ once its
// pass has ended (what tracePixel does when its path ends).
template <class Scenes>
__global__ __launch_bounds__(256) void kWfToXYZ(Scenes S, WfState w) {
    const uint32_t n = w.active * w.npix;
    PERSISTENT_LOOP(p, n) {
        const uint32_t fl = w.flags[p];
        stS(w.rad, p, laneXYZ(ldS(w.rad, p), pathLanes(S.view(sceneOf(fl)), fl)));
    }
}

// Training samples from this pass's records: incident radiance along the
// guided direction = (final radiance - radiance before that direction's
// contributions) / throughput; target = luminance of (BSDF * cos) x that,
// weight = target / sampling pdf.  With radianceTarget the target is the
// mean of the incident radiance alone (no BSDF factor).  Records whose path ended before the next
// vertex (no snapshot) or with zero target are skipped; at most cap samples
// are written (the counter still counts them all).
__global__ __launch_bounds__(256) void kWfRecordsToSamples(WfState w, GuideSample* out,
                                                           uint32_t cap,
                                                           uint32_t radianceTarget) {
    const uint32_t n = min(w.counters[kCRec], w.recCap);
    PERSISTENT_LOOP(r, n) {
        const float4* R = w.rec + size_t(r) * kRecF4;
        float4 snap = R[6];
        bool ok = snap.w == 1.f;
        float weight = 0.f;
        if (ok) {
            uint32_t p = __float_as_uint(R[3].w);
            float4 e = w.rad[p], f = R[4], bw = R[5];
            float li[3] = {bw.x > 0.f ? (e.x - snap.x) / bw.x : 0.f,
                           bw.y > 0.f ? (e.y - snap.y) / bw.y : 0.f,
                           bw.z > 0.f ? (e.z - snap.z) / bw.z : 0.f};
            float y = radianceTarget
                          ? (li[0] + li[1] + li[2]) / 3.f
                          : (f.x * li[0] + f.y * li[1] + f.z * li[2]) / 3.f;
            weight = y / R[0].w;
            ok = weight > 0.f && isfinite(weight);
        }
        uint32_t slot = wavePush(&w.counters[kCTrain], ok);
        if (ok && slot < cap) {
            GuideSample& g = out[slot];
            float4 a = R[0], b = R[1], c = R[2], d = R[3];
            g.pos[0] = a.x; g.pos[1] = a.y; g.pos[2] = a.z; g.pos[3] = 0.f;
            g.nrm[0] = b.x; g.nrm[1] = b.y; g.nrm[2] = b.z; g.nrm[3] = 0.f;
            g.wo[0] = c.x; g.wo[1] = c.y; g.wo[2] = c.z; g.wo[3] = 0.f;
            g.wi[0] = d.x; g.wi[1] = d.y; g.wi[2] = d.z; g.wi[3] = weight;
        }
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
template <int SM, class Scenes>
static void launchExtend(bool count, dim3 g, dim3 b, size_t lds, hipStream_t st,
                         const Scenes& S, const WfState& w, uint32_t cur) {
    if (count)
        hipLaunchKernelGGL((kWfExtend<true, SM, Scenes>), g, b, lds, st, S, w, cur);
    else
        hipLaunchKernelGGL((kWfExtend<false, SM, Scenes>), g, b, lds, st, S, w, cur);
}

// One wavefront pass over the scene(s) of S.  stackCap: the deepest tree of
// them; cam: the camera of a single scene (a batch's come from its table).
// Batched passes trace one ray per thread (no lane refill: the persistent
// traversal kernels hold one scene).
template <class Scenes>
static hipError_t wfPassImpl(const Scenes& S, uint32_t stackCap, const CameraView& cam,
                             const IntegratorView& cfg, uint32_t si, uint32_t active, WfState& w,
                             const WfLaunch& L, bool countWork, hipStream_t st,
                             const GuideView* guide, float alpha, GuideSample* samples,
                             uint32_t sampleCap, uint32_t guideMask, uint32_t guideDepth,
                             uint32_t guideTarget, bool spectral) {
    if (active == 0 || active > w.slots) return hipErrorInvalidValue;
    if (spectral && (guide || countWork)) return hipErrorInvalidValue;
    if (Scenes::kBatch && (guide || countWork)) return hipErrorInvalidValue;
    w.active = active;
    const uint32_t TB = L.traceBlock;
    const size_t lds = stackLdsBytes(L.privStack, stackCap, TB);
    const dim3 b256(256), bT(TB);
    const bool guided = guide != nullptr;
    const bool pt = L.persistent != 0 && !Scenes::kBatch;
    auto extend = [&](uint32_t cur, bool count) {
        if (pt && !count)
            hipLaunchKernelGGL(kWfReset, dim3(1), dim3(64), 0, st, w.counters, 1u << kCPtExt);
        withStack(L.privStack, [&](auto sm) {
            constexpr int SM = decltype(sm)::value;
            if constexpr (!Scenes::kBatch) {
                if (pt && !count) {
                    hipLaunchKernelGGL((kWfExtendPT<SM>), dim3(L.traceGrid), bT, lds, st,
                                       S.s, w, cur, L.refill);
                    return;
                }
            }
            launchExtend<SM>(count, dim3(L.traceGrid), bT, lds, st, S, w, cur);
        });
    };
    auto shadow = [&]() {
        if (pt)
            hipLaunchKernelGGL(kWfReset, dim3(1), dim3(64), 0, st, w.counters, 1u << kCPtShadow);
        withStack(L.privStack, [&](auto sm) {
            constexpr int SM = decltype(sm)::value;
            if constexpr (!Scenes::kBatch) {
                if (pt) {
                    hipLaunchKernelGGL((kWfShadowPT<SM>), dim3(L.traceGrid), bT, lds, st,
                                       S.s, w, L.refill);
                    return;
                }
            }
            hipLaunchKernelGGL((kWfShadow<SM, Scenes>), dim3(L.traceGrid), bT, lds, st, S, w);
        });
    };
    hipLaunchKernelGGL(kWfReset, dim3(1), dim3(64), 0, st, w.counters,
                       (1u << kCExt1) | (1u << kCMis0) | (1u << kCMis1) |
                           (1u << kCShadow) | (1u << kCCont) | (1u << kCRec) |
                           (1u << kCTrain));
    if (spectral)
        hipLaunchKernelGGL((kWfGenerate<true, Scenes>), dim3(L.shadeGrid), b256, 0, st, w, cam,
                           si, S);
    else
        hipLaunchKernelGGL((kWfGenerate<false, Scenes>), dim3(L.shadeGrid), b256, 0, st, w, cam,
                           si, S);
    auto misResolve = [&](uint32_t c) {
        if (spectral)
            hipLaunchKernelGGL((kWfMisResolve<true, Scenes>), dim3(L.shadeGrid), b256, 0, st, S,
                               w, c);
        else
            hipLaunchKernelGGL((kWfMisResolve<false, Scenes>), dim3(L.shadeGrid), b256, 0, st, S,
                               w, c);
    };
    uint32_t cur = 0;
    // Bounces run as wavefront launches up to finishDepth; the paths still
    // in flight then finish in one per-path launch.
    const bool finish = L.finishDepth > 0u && L.finishDepth < cfg.maxDepth && !countWork;
    const uint32_t dEnd = finish ? L.finishDepth : cfg.maxDepth;
    for (uint32_t d = 0; d < dEnd; ++d) {
        const uint32_t nxt = cur ^ 1u;
        extend(cur, countWork);
        misResolve(cur);
        hipLaunchKernelGGL(kWfReset, dim3(1), dim3(64), 0, st, w.counters,
                           (1u << (kCExt0 + nxt)) | (1u << (kCMis0 + nxt)) |
                               (1u << kCShadow) | (1u << kCCont));
        if constexpr (!Scenes::kBatch) {
            if (guided)
                hipLaunchKernelGGL((kWfShadeHit<true, false, Scenes>), dim3(L.shadeGrid), b256, 0,
                                   st, S, w, cur);
        }
        if (!guided && spectral)
            hipLaunchKernelGGL((kWfShadeHit<false, true, Scenes>), dim3(L.shadeGrid), b256, 0, st,
                               S, w, cur);
        else if (!guided)
            hipLaunchKernelGGL((kWfShadeHit<false, false, Scenes>), dim3(L.shadeGrid), b256, 0,
                               st, S, w, cur);
        shadow();
        // Past guideDepth the guided render continues like the plain one (no
        // inference); records of earlier vertices still get their snapshots.
        const bool guideHere = guided && guideMask != 0u && (guideDepth == 0u || d < guideDepth);
        const uint32_t gm = guideMask;
#define CRH_WFC(MIS, GUIDED, SPEC, GM)                                                      \
    hipLaunchKernelGGL((kWfShadeCont<MIS, GUIDED, SPEC, Scenes>), dim3(L.shadeGrid), b256, 0, st, \
                       S, w, cfg, cur, d, alpha, GM)
        if (guideHere) {
            if constexpr (!Scenes::kBatch) {
                hipError_t e = launchGuideInfer(*guide, w.qCont, w.counters + kCCont, 0u, w.gPos,
                                                w.gNrm, w.gWo, w.gMix, nullptr, L.shadeGrid, st);
                if (e != hipSuccess) return e;
                if (cfg.useMIS) CRH_WFC(true, true, false, gm);
                else CRH_WFC(false, true, false, gm);
            }
        } else if (spectral) {
            if (cfg.useMIS) CRH_WFC(true, false, true, 0u);
            else CRH_WFC(false, false, true, 0u);
        } else {
            if (cfg.useMIS) CRH_WFC(true, false, false, 0u);
            else CRH_WFC(false, false, false, 0u);
        }
#undef CRH_WFC
        cur = nxt;
    }
    if (finish) {
        hipLaunchKernelGGL(kWfReset, dim3(1), dim3(64), 0, st, w.counters, 1u << kCPtExt);
        WfFinishArgs fa{};
        fa.w = w;
        fa.cfg = cfg;
        fa.cur = cur;
        fa.d0 = dEnd;
        fa.grid = L.finishGrid;
        fa.lds = stackLdsBytes(L.privStack, stackCap, 256);
        fa.waves = L.finishWaves;
        fa.spectral = spectral;
        fa.st = st;
        if constexpr (Scenes::kBatch) {
            fa.batch = true;
            fa.tab = S.tab;
            fa.pixPrefix = S.pixPrefix;
            fa.nActive = S.nActive;
        } else {
            fa.s = S.s;
        }
        hipError_t e = hipSuccess;
        withStack(L.privStack, [&](auto sm) { e = launchWfFinishSM<decltype(sm)::value>(fa); });
        if (e != hipSuccess) return e;
    } else {
        // MIS rays spawned at the last bounce: trace and resolve them.
        hipLaunchKernelGGL(kWfReset, dim3(1), dim3(64), 0, st, w.counters,
                           1u << (kCExt0 + cur));
        extend(cur, false);
        misResolve(cur);
    }
    if (spectral)
        hipLaunchKernelGGL((kWfToXYZ<Scenes>), dim3(L.shadeGrid), b256, 0, st, S, w);
    if (guided && samples)
        hipLaunchKernelGGL(kWfRecordsToSamples, dim3(L.shadeGrid), b256, 0, st, w, samples,
                           sampleCap, guideTarget);
    return hipGetLastError();
}

hipError_t wavefrontPass(const SceneView& s, const CameraView& cam,
                         const IntegratorView& cfg, uint32_t si, uint32_t active, WfState& w,
                         const WfLaunch& L, bool countWork, hipStream_t st,
                         const GuideView* guide, float alpha, GuideSample* samples,
                         uint32_t sampleCap, uint32_t guideMask, uint32_t guideDepth,
                         uint32_t guideTarget, bool spectral) {
    if (spectral && !s.spec) return hipErrorInvalidValue;
    return wfPassImpl(OneScene{s}, s.stackCap, cam, cfg, si, active, w, L, countWork, st, guide,
                      alpha, samples, sampleCap, guideMask, guideDepth, guideTarget, spectral);
}

hipError_t wavefrontBatchPass(const BatchScene* tab, const uint32_t* pixPrefix, uint32_t nActive,
                              uint32_t stackCap, const IntegratorView& cfg, uint32_t si,
                              uint32_t active, WfState& w, const WfLaunch& L, hipStream_t st,
                              bool spectral) {
    if (nActive == 0 || nActive > kSceneMask + 1u) return hipErrorInvalidValue;
    return wfPassImpl(SceneTable{tab, pixPrefix, nActive}, stackCap, CameraView{}, cfg, si,
                      active, w, L, false, st, nullptr, 0.5f, nullptr, 0u, 3u, 0u, 0u, spectral);
}

}  // namespace crh
