// gfx950 kernels: batch closest-hit / any-hit queries, the plain megakernel
// integrator (one path per thread per sample index) and the deterministic
// film gather.  Host launchers at the bottom are called by api.cpp.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "integrator.h"

namespace crh {

// ---------------------------------------------------------------------------
// Ray batch queries
// ---------------------------------------------------------------------------
template <bool COUNT, bool PRIV>
__global__ __launch_bounds__(kTraceBlock) void kTraceClosest(
    SceneView s, const crh_ray* __restrict__ rays, uint64_t n,
    crh_hit* __restrict__ hits, uint32_t* __restrict__ work) {
    extern __shared__ uint2 lds[];
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4* rp = reinterpret_cast<const float4*>(rays + i);
    float4 a = rp[0], b = rp[1];
    Ray ray{mk3(a.x, a.y, a.z), a.w, mk3(b.x, b.y, b.z), b.w};
    Hit h;
    WorkCount wc{0, 0};
    auto stk = makeStack<PRIV>(lds);
    traverse<false, COUNT>(s, ray, h, stk, wc);
    reinterpret_cast<uint4*>(hits)[i] =
        make_uint4(h.prim, __float_as_uint(h.t), __float_as_uint(h.u),
                   __float_as_uint(h.v));
    if (COUNT) {
        work[2 * i] = wc.nodes;
        work[2 * i + 1] = wc.tris;
    }
}

template <bool PRIV>
__global__ __launch_bounds__(kTraceBlock) void kTraceAny(
    SceneView s, const crh_ray* __restrict__ rays, uint64_t n,
    uint8_t* __restrict__ occ) {
    extern __shared__ uint2 lds[];
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4* rp = reinterpret_cast<const float4*>(rays + i);
    float4 a = rp[0], b = rp[1];
    Ray ray{mk3(a.x, a.y, a.z), a.w, mk3(b.x, b.y, b.z), b.w};
    Hit h;
    WorkCount wc{0, 0};
    auto stk = makeStack<PRIV>(lds);
    occ[i] = traverse<true, false>(s, ray, h, stk, wc) ? 1 : 0;
}

// Persistent variants: lanes refill with new rays as theirs finish.
struct BatchSrc {
    const crh_ray* rays;
    crh_hit* hits;
    uint8_t* occ;
    CRH_D Ray load(uint32_t i) const {
        const float4* rp = reinterpret_cast<const float4*>(rays + i);
        float4 a = rp[0], b = rp[1];
        return Ray{mk3(a.x, a.y, a.z), a.w, mk3(b.x, b.y, b.z), b.w};
    }
    CRH_D void store(uint32_t i, bool isHit, const Hit& h) const {
        if (hits)
            reinterpret_cast<uint4*>(hits)[i] =
                make_uint4(h.prim, __float_as_uint(h.t), __float_as_uint(h.u), __float_as_uint(h.v));
        if (occ) occ[i] = isHit ? 1 : 0;
    }
};

__global__ __launch_bounds__(64) void kResetCounter(uint32_t* c) {
    if (threadIdx.x == 0) *c = 0u;
}

template <bool ANY, bool PRIV>
__global__ __launch_bounds__(kTraceBlock) void kTracePersistent(
    SceneView s, BatchSrc src, uint32_t n, uint32_t* __restrict__ counter, uint32_t refill) {
    extern __shared__ uint2 lds[];
    auto stk = makeStack<PRIV>(lds);
    WorkCount wc{0, 0};
    persistentTrace<ANY, false>(s, src, counter, n, refill, stk, wc);
}

// ---------------------------------------------------------------------------
// Megakernel integrator (Core/Integrator.cu:13-227)
// ---------------------------------------------------------------------------
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
                              const BSDF& bsdf, Sampler& smp) {
    LightGeom g = lightGeom(c.s, it, smp);
    if (!g.valid) return noLight();
    if (occluded(c, g.shadow)) return noLight();
    return lightFinish(c.s, g, bsdf, it);
}

// evalMIS (Core/Integrator.cu:74-140)
template <bool COUNT, class Stack>
CRH_D Spec evalMIS(PathCtx<COUNT, Stack>& c, const Isect& it, const BSDF& bsdf,
                   Sampler& smp) {
    LightSample ls = sampleLight(c, it, bsdf, smp);
    if (ls.pdf <= 0.0f) return sp(0.f);
    Spec value = misLightPart(ls, bsdf, it);
    Ray r;
    Spec bw;
    float bp;
    if (misBsdfRay(bsdf, smp, it, r, bw, bp)) {
        Hit h;
        bool isHit = closest(c, r, h);
        misBsdfPart(c.s, it.pos, r, isHit, h, ls.id, bw, bp, value);
    }
    return value / ls.selPdf;
}

CRH_D void tileCoord(uint32_t tilesX, uint32_t& x, uint32_t& y) {
    // 16x16 pixel tile per block; each wave covers an 8x8 quarter.
    uint32_t tile = blockIdx.x;
    uint32_t tx = tile % tilesX, ty = tile / tilesX;
    uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    x = tx * 16u + (wave & 1u) * 8u + (lane & 7u);
    y = ty * 16u + (wave >> 1) * 8u + (lane >> 3);
}

// pathTraceKernel (Core/Integrator.cu:142-227): radiance of sample index
// `si` for every pixel into L (float4 per pixel, w unused) and the jittered
// film position into P.
template <bool MIS, bool COUNT, bool PRIV>
__global__ __launch_bounds__(kRenderBlock) void kRenderPass(
    SceneView s, CameraView cam, IntegratorView cfg, uint32_t si,
    uint32_t tilesX, float4* __restrict__ L, float2* __restrict__ P,
    unsigned long long* __restrict__ counters) {
    extern __shared__ uint2 lds[];
    uint32_t x, y;
    tileCoord(tilesX, x, y);
    uint32_t rays = 0, nodes = 0, tris = 0;
    if (x < cam.width && y < cam.height) {
        auto stk = makeStack<PRIV>(lds);
        PathCtx<COUNT, typename StackOf<PRIV>::T> c{s, stk, 0, {0, 0}};
        Sampler smp(x, y, si);
        F2 j = smp.next2D();
        F2 pix{float(x) + j.x, float(y) + j.y};
        Ray ray = cameraRay(cam, pix);
        Spec rad = sp(0.f), beta = sp(1.f);
        for (uint32_t depth = 0; depth < cfg.maxDepth; ++depth) {
            bool terminate = false;
            Hit h;
            if (closest(c, ray, h)) {
                Isect it = makeIsect(s, ray, h);
                const GMaterial& m = s.mats[s.primMat[h.prim]];
                if (m.flags & 1u) {
                    rad = rad + sp3(m.emission) * beta;
                    break;
                }
                BSDF bsdf = makeBSDF(m, it.frame);
                if (MIS) {
                    rad = rad + evalMIS(c, it, bsdf, smp) * beta;
                } else {
                    LightSample ls = sampleLight(c, it, bsdf, smp);
                    if (ls.pdf > 0.0f)
                        rad = rad + ls.weight * beta / ls.pdf / ls.selPdf;
                }
                F3 wi;
                float pdf = 0.0f;
                Spec f = bsdfSampleEval(bsdf, smp, it.view, wi, pdf);
                if (pdf <= 0.0f) break;
                beta = beta * (f / pdf);
                ray.o = offsetRayOrigin(it.pos, oriented(it));
                ray.d = wi;
            } else {
                rad = rad + sp3(s.env) * beta;
                terminate = true;
            }
            if (smp.next1D() < cfg.rrThreshold) terminate = true;
            if (terminate) break;
            if (!roulette(cfg, depth, smp, beta)) break;
        }
        size_t id = size_t(y) * cam.width + x;
        L[id] = make_float4(rad.r, rad.g, rad.b, 0.f);
        P[id] = make_float2(pix.x, pix.y);
        rays = c.rays;
        nodes = c.wc.nodes;
        tris = c.wc.tris;
    }
    // Wave-level reductions, one atomic per wave and counter:
    // counters = {rays, node visits, triangle tests}.
    unsigned long long r64 = rays, n64 = nodes, t64 = tris;
    for (int off = 32; off > 0; off >>= 1) {
        r64 += __shfl_down(r64, off, 64);
        if (COUNT) {
            n64 += __shfl_down(n64, off, 64);
            t64 += __shfl_down(t64, off, 64);
        }
    }
    if ((threadIdx.x & 63u) == 0) {
        if (r64) atomicAdd(counters, r64);
        if (COUNT) {
            atomicAdd(counters + 1, n64);
            atomicAdd(counters + 2, t64);
        }
    }
}

// Deterministic film update for one pass.  Each pixel gathers the tent-filter
// contributions of the 3x3 neighbouring samples in row-major sample order,
// which is the order a sequential SensorData::addSample loop over the pass
// (Core/Sensor.cu:19-43) delivers them; samples with NaN/Inf are dropped and
// zero weights skipped (they add exactly nothing).
__global__ __launch_bounds__(256) void kFilmGather(
    const float4* __restrict__ L, const float2* __restrict__ P, uint32_t W,
    uint32_t H, float* __restrict__ sum, float* __restrict__ wsum) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= W * H) return;
    int x = int(i % W), y = int(i / W);
    float s0 = sum[3 * size_t(i)], s1 = sum[3 * size_t(i) + 1],
          s2 = sum[3 * size_t(i) + 2], ws = wsum[i];
    for (int dy = -1; dy <= 1; ++dy) {
        int sy = y + dy;
        if (sy < 0 || sy >= int(H)) continue;
        for (int dx = -1; dx <= 1; ++dx) {
            int sx = x + dx;
            if (sx < 0 || sx >= int(W)) continue;
            size_t sidx = size_t(sy) * W + sx;
            float4 l = L[sidx];
            if (isnan(l.x) || isnan(l.y) || isnan(l.z) || isinf(l.x) ||
                isinf(l.y) || isinf(l.z))
                continue;
            float2 p = P[sidx];
            float ddx = p.x - (float(x) + 0.5f);
            float ddy = p.y - (float(y) + 0.5f);
            float w = smax(0.0f, 1.0f - fabsf(ddx)) * smax(0.0f, 1.0f - fabsf(ddy));
            if (w == 0.0f) continue;
            s0 += l.x * w;
            s1 += l.y * w;
            s2 += l.z * w;
            ws += w;
        }
    }
    sum[3 * size_t(i)] = s0;
    sum[3 * size_t(i) + 1] = s1;
    sum[3 * size_t(i) + 2] = s2;
    wsum[i] = ws;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static size_t stackLds(const SceneView& s, uint32_t priv, uint32_t block) {
    return priv ? 0 : size_t(s.stackCap) * block * sizeof(uint2);
}

template <bool ANY>
static hipError_t launchPersistent(const SceneView& s, const BatchSrc& src, uint32_t n,
                                   uint32_t* counter, const Tuning& t, hipStream_t st) {
    const uint32_t B = t.traceBlock;
    size_t lds = stackLds(s, t.privStack, B);
    hipLaunchKernelGGL(kResetCounter, dim3(1), dim3(64), 0, st, counter);
    dim3 grid(t.ptGrid), block(B);
    if (t.privStack)
        hipLaunchKernelGGL((kTracePersistent<ANY, true>), grid, block, lds, st, s, src, n, counter,
                           t.refill);
    else
        hipLaunchKernelGGL((kTracePersistent<ANY, false>), grid, block, lds, st, s, src, n,
                           counter, t.refill);
    return hipGetLastError();
}

hipError_t launchTraceClosest(const SceneView& s, const crh_ray* rays,
                              uint64_t n, crh_hit* hits, uint32_t* work,
                              const Tuning& t, uint32_t* counter, hipStream_t st) {
    if (n == 0) return hipSuccess;
    if (t.persistent && !work && counter && n < (uint64_t(1) << 32))
        return launchPersistent<false>(s, BatchSrc{rays, hits, nullptr}, uint32_t(n), counter, t,
                                       st);
    const uint32_t B = t.traceBlock;
    size_t lds = stackLds(s, t.privStack, B);
    dim3 grid(uint32_t((n + B - 1) / B)), block(B);
    if (work && t.privStack)
        hipLaunchKernelGGL((kTraceClosest<true, true>), grid, block, lds, st, s, rays, n, hits, work);
    else if (work)
        hipLaunchKernelGGL((kTraceClosest<true, false>), grid, block, lds, st, s, rays, n, hits, work);
    else if (t.privStack)
        hipLaunchKernelGGL((kTraceClosest<false, true>), grid, block, lds, st, s, rays, n, hits, work);
    else
        hipLaunchKernelGGL((kTraceClosest<false, false>), grid, block, lds, st, s, rays, n, hits, work);
    return hipGetLastError();
}

hipError_t launchTraceAny(const SceneView& s, const crh_ray* rays, uint64_t n,
                          uint8_t* occ, const Tuning& t, uint32_t* counter, hipStream_t st) {
    if (n == 0) return hipSuccess;
    if (t.persistent && counter && n < (uint64_t(1) << 32))
        return launchPersistent<true>(s, BatchSrc{rays, nullptr, occ}, uint32_t(n), counter, t, st);
    const uint32_t B = t.traceBlock;
    size_t lds = stackLds(s, t.privStack, B);
    dim3 grid(uint32_t((n + B - 1) / B)), block(B);
    if (t.privStack)
        hipLaunchKernelGGL((kTraceAny<true>), grid, block, lds, st, s, rays, n, occ);
    else
        hipLaunchKernelGGL((kTraceAny<false>), grid, block, lds, st, s, rays, n, occ);
    return hipGetLastError();
}

template <bool MIS, bool COUNT, bool PRIV>
static void renderLaunch(dim3 grid, size_t lds, hipStream_t st, const SceneView& s,
                         const CameraView& cam, const IntegratorView& cfg, uint32_t si,
                         uint32_t tilesX, float4* L, float2* P,
                         unsigned long long* counters) {
    hipLaunchKernelGGL((kRenderPass<MIS, COUNT, PRIV>), grid, dim3(kRenderBlock), lds, st, s,
                       cam, cfg, si, tilesX, L, P, counters);
}

hipError_t launchRenderPass(const SceneView& s, const CameraView& cam,
                            const IntegratorView& cfg, uint32_t si, float4* L,
                            float2* P, unsigned long long* counters,
                            bool countWork, const Tuning& t, hipStream_t st) {
    uint32_t tilesX = (cam.width + 15u) / 16u, tilesY = (cam.height + 15u) / 16u;
    size_t lds = stackLds(s, t.renderPriv, kRenderBlock);
    dim3 grid(tilesX * tilesY);
    const int sel = (cfg.useMIS ? 4 : 0) | (countWork ? 2 : 0) | (t.renderPriv ? 1 : 0);
    switch (sel) {
        case 0: renderLaunch<false, false, false>(grid, lds, st, s, cam, cfg, si, tilesX, L, P, counters); break;
        case 1: renderLaunch<false, false, true>(grid, lds, st, s, cam, cfg, si, tilesX, L, P, counters); break;
        case 2: renderLaunch<false, true, false>(grid, lds, st, s, cam, cfg, si, tilesX, L, P, counters); break;
        case 3: renderLaunch<false, true, true>(grid, lds, st, s, cam, cfg, si, tilesX, L, P, counters); break;
        case 4: renderLaunch<true, false, false>(grid, lds, st, s, cam, cfg, si, tilesX, L, P, counters); break;
        case 5: renderLaunch<true, false, true>(grid, lds, st, s, cam, cfg, si, tilesX, L, P, counters); break;
        case 6: renderLaunch<true, true, false>(grid, lds, st, s, cam, cfg, si, tilesX, L, P, counters); break;
        default: renderLaunch<true, true, true>(grid, lds, st, s, cam, cfg, si, tilesX, L, P, counters); break;
    }
    return hipGetLastError();
}

hipError_t launchFilmGather(const float4* L, const float2* P, uint32_t W,
                            uint32_t H, float* sum, float* wsum,
                            hipStream_t st) {
    uint32_t n = W * H;
    hipLaunchKernelGGL(kFilmGather, dim3((n + 255u) / 256u), dim3(256), 0, st,
                       L, P, W, H, sum, wsum);
    return hipGetLastError();
}

}  // namespace crh
