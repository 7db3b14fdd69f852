// gfx950 kernels: batch closest-hit / any-hit queries and the deterministic
// film gather; host launchers (called by api.cpp) at the bottom.  The
// megakernels themselves (megakernel.h) are instantiated in mk_inst.hip,
// one translation unit per kernel family and stack mode.
#include <hip/hip_runtime.h>

#include "megakernel.h"

namespace crh {

// ---------------------------------------------------------------------------
// Ray batch queries
// ---------------------------------------------------------------------------
template <bool COUNT, int SM>
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
    auto stk = makeStack<SM>(lds);
    traverse<false, COUNT>(s, ray, h, stk, wc);
    reinterpret_cast<uint4*>(hits)[i] =
        make_uint4(h.prim, __float_as_uint(h.t), __float_as_uint(h.u),
                   __float_as_uint(h.v));
    if (COUNT) {
        work[2 * i] = wc.nodes;
        work[2 * i + 1] = wc.tris;
    }
}

template <int SM>
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
    auto stk = makeStack<SM>(lds);
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

template <bool ANY, int SM>
__global__ __launch_bounds__(kTraceBlock) void kTracePersistent(
    SceneView s, BatchSrc src, uint32_t n, uint32_t* __restrict__ counter, uint32_t refill,
    uint32_t tris) {
    extern __shared__ uint2 lds[];
    auto stk = makeStack<SM>(lds);
    WorkCount wc{0, 0};
    persistentTrace<ANY, false>(s, src, counter, n, refill, stk, wc, tris);
}

__global__ void kResetNext(uint32_t* __restrict__ next) {
    if (threadIdx.x == 0) *next = 0u;
}

// Deterministic film update for one pass.  Each pixel gathers the tent-filter
// contributions of the 3x3 neighbouring samples in row-major sample order,
// which is the order a sequential SensorData::addSample loop over the pass
// (Core/Sensor.cu:19-43) delivers them; samples with NaN/Inf are dropped and
// zero weights skipped (they add exactly nothing).
CRH_D void gatherAdd(const float4* __restrict__ L, const float2* __restrict__ P, uint32_t W,
                     uint32_t H, int x, int y, float& s0, float& s1, float& s2, float& ws) {
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
}
CRH_D void gatherPixel(const float4* __restrict__ L, const float2* __restrict__ P,
                       uint32_t W, uint32_t H, float* __restrict__ sum,
                       float* __restrict__ wsum, uint32_t i) {
    float s0 = sum[3 * size_t(i)], s1 = sum[3 * size_t(i) + 1],
          s2 = sum[3 * size_t(i) + 2], ws = wsum[i];
    gatherAdd(L, P, W, H, int(i % W), int(i / W), s0, s1, s2, ws);
    sum[3 * size_t(i)] = s0;
    sum[3 * size_t(i) + 1] = s1;
    sum[3 * size_t(i) + 2] = s2;
    wsum[i] = ws;
}

// `slots` consecutive sample slots (slot k at L + k * W * H), one after the
// other, with the pixel's sums kept in registers in between: the same
// additions in the same order as one gather per slot.
__global__ __launch_bounds__(256) void kFilmGather(
    const float4* __restrict__ L, const float2* __restrict__ P, uint32_t W,
    uint32_t H, uint32_t slots, float* __restrict__ sum, float* __restrict__ wsum) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= W * H) return;
    const size_t npix = size_t(W) * H;
    float s0 = sum[3 * size_t(i)], s1 = sum[3 * size_t(i) + 1],
          s2 = sum[3 * size_t(i) + 2], ws = wsum[i];
    for (uint32_t k = 0; k < slots; ++k)
        gatherAdd(L + k * npix, P + k * npix, W, H, int(i % W), int(i / W), s0, s1, s2, ws);
    sum[3 * size_t(i)] = s0;
    sum[3 * size_t(i) + 1] = s1;
    sum[3 * size_t(i) + 2] = s2;
    wsum[i] = ws;
}

__global__ void kBatchReset(uint32_t* __restrict__ next) {
    if (threadIdx.x < kBatchMaxChunks) next[threadIdx.x] = 0u;
}

// Film update of every active scene with the samples of slot `slot`:
// pixPrefix[j] is scene j's first pixel.  The samples are the scenes' own
// buffers (slot k of scene j at b.L + k * w * h), or with Ls set a pass
// buffer holding every active scene (slot k of scene j at
// Ls + k * stride + pixPrefix[j], as the batched wavefront lays them out).
__global__ __launch_bounds__(256) void kBatchGather(const BatchScene* __restrict__ tab,
                                                    const uint32_t* __restrict__ pixPrefix,
                                                    uint32_t nActive, uint32_t slot,
                                                    const float4* __restrict__ Ls,
                                                    const float2* __restrict__ Ps,
                                                    uint32_t stride) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= pixPrefix[nActive]) return;
    const uint32_t j = findScene(pixPrefix, nActive, i);
    const BatchScene& b = tab[j];
    const float4* L;
    const float2* P;
    if (Ls) {
        const size_t off = size_t(slot) * stride + pixPrefix[j];
        L = Ls + off;
        P = Ps + off;
    } else {
        const size_t off = size_t(slot) * b.cam.width * b.cam.height;
        L = reinterpret_cast<const float4*>(b.L) + off;
        P = reinterpret_cast<const float2*>(b.P) + off;
    }
    gatherPixel(L, P, b.cam.width, b.cam.height, b.sum, b.wsum, i - pixPrefix[j]);
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
size_t stackLdsBytes(uint32_t mode, uint32_t stackCap, uint32_t block) {
    const uint32_t entries = mode == 1 ? 0u : (mode == 2 ? uint32_t(kShortStackLds) : stackCap);
    return size_t(entries) * block * sizeof(uint2);
}

template <bool ANY>
static hipError_t launchPersistent(const SceneView& s, const BatchSrc& src, uint32_t n,
                                   uint32_t* counter, const Tuning& t, hipStream_t st) {
    const uint32_t B = t.traceBlock;
    const size_t lds = stackLdsBytes(t.privStack, s.stackCap, B);
    hipLaunchKernelGGL(kResetCounter, dim3(1), dim3(64), 0, st, counter);
    const dim3 grid(t.ptGrid), block(B);
    withStack(t.privStack, [&](auto sm) {
        hipLaunchKernelGGL((kTracePersistent<ANY, decltype(sm)::value>), grid, block, lds, st, s,
                           src, n, counter, t.refill, t.traceTris);
    });
    return hipGetLastError();
}

hipError_t launchTraceClosest(const SceneView& s, const crh_ray* rays,
                              uint64_t n, crh_hit* hits, uint32_t* work,
                              const Tuning& t, uint32_t* counter, hipStream_t st) {
    if (n == 0) return hipSuccess;
    // The persistent kernels' dispenser counts in 32 bits and may run up to
    // (waves of the launch) x (chunk <= 512) past n before every wave sees
    // the end: n <= 2^31 keeps that far below 2^32.
    if (t.persistent && !work && counter && n <= (uint64_t(1) << 31))
        return launchPersistent<false>(s, BatchSrc{rays, hits, nullptr}, uint32_t(n), counter, t,
                                       st);
    const uint32_t B = t.traceBlock;
    const size_t lds = stackLdsBytes(t.privStack, s.stackCap, B);
    const dim3 grid(uint32_t((n + B - 1) / B)), block(B);
    withStack(t.privStack, [&](auto sm) {
        constexpr int SM = decltype(sm)::value;
        if (work)
            hipLaunchKernelGGL((kTraceClosest<true, SM>), grid, block, lds, st, s, rays, n, hits,
                               work);
        else
            hipLaunchKernelGGL((kTraceClosest<false, SM>), grid, block, lds, st, s, rays, n, hits,
                               work);
    });
    return hipGetLastError();
}

hipError_t launchTraceAny(const SceneView& s, const crh_ray* rays, uint64_t n,
                          uint8_t* occ, const Tuning& t, uint32_t* counter, hipStream_t st) {
    if (n == 0) return hipSuccess;
    if (t.persistent && counter && n <= (uint64_t(1) << 31))  // as launchTraceClosest
        return launchPersistent<true>(s, BatchSrc{rays, nullptr, occ}, uint32_t(n), counter, t, st);
    const uint32_t B = t.traceBlock;
    const size_t lds = stackLdsBytes(t.privStack, s.stackCap, B);
    const dim3 grid(uint32_t((n + B - 1) / B)), block(B);
    withStack(t.privStack, [&](auto sm) {
        hipLaunchKernelGGL((kTraceAny<decltype(sm)::value>), grid, block, lds, st, s, rays, n,
                           occ);
    });
    return hipGetLastError();
}

hipError_t launchRenderPass(const SceneView& s, const CameraView& cam,
                            const IntegratorView& cfg, uint32_t si, uint32_t active, float4* L,
                            float2* P, unsigned long long* counters,
                            bool countWork, const Tuning& t, hipStream_t st, bool spectral,
                            uint32_t* next) {
    if (spectral && (!s.spec || countWork)) return hipErrorInvalidValue;
    const uint32_t tilesX = (cam.width + 15u) / 16u, tilesY = (cam.height + 15u) / 16u;
    const uint32_t tiles = tilesX * tilesY;
    if (uint64_t(tiles) * active >= (uint64_t(1) << 23)) return hipErrorInvalidValue;
    MkArgs a{s, cam, cfg, si, tilesX, tiles, t.xcdBand, 0u, next, L, P, counters,
             dim3(tiles * active), stackLdsBytes(t.renderPriv, s.stackCap, kRenderBlock),
             t.renderWaves, cfg.useMIS != 0u, countWork, spectral, st};
    if (t.renderRegen && !countWork && next) {
        a.total = tiles * active * 256u;
        a.grid = dim3(std::min<uint32_t>(tiles * active, t.renderGrid));
        hipLaunchKernelGGL(kResetNext, dim3(1), dim3(64), 0, st, next);
        hipError_t e = hipSuccess;
        withStack(t.renderPriv, [&](auto sm) { e = launchRegenSM<decltype(sm)::value>(a); });
        return e != hipSuccess ? e : hipGetLastError();
    }
    hipError_t e = hipSuccess;
    withStack(t.renderPriv, [&](auto sm) { e = launchRenderSM<decltype(sm)::value>(a); });
    return e != hipSuccess ? e : hipGetLastError();
}

hipError_t launchBatchPass(const BatchScene* tab, const uint32_t* tilePrefix,
                           const uint32_t* pixPrefix, uint32_t nActive, uint32_t items,
                           uint32_t pixels, const IntegratorView& cfg, uint32_t si,
                           uint32_t slots, uint32_t stackCap, const Tuning& t, uint32_t* next,
                           unsigned long long* counters, hipStream_t st, bool spectral) {
    if (nActive == 0 || items == 0 || slots == 0) return hipSuccess;
    if (uint64_t(items) * slots >= (uint64_t(1) << 32)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kBatchReset, dim3(1), dim3(64), 0, st, next);
    const uint32_t total = items * slots;
    const BatchArgs a{tab, tilePrefix, nActive, cfg, si, slots, t.batchTileW, t.batchChunks,
                      next, counters, dim3(total < t.batchGrid ? total : t.batchGrid),
                      stackLdsBytes(t.batchPriv, stackCap, kRenderBlock), t.batchWaves,
                      cfg.useMIS != 0u, spectral, st};
    hipError_t e = hipSuccess;
    withStack(t.batchPriv, [&](auto sm) { e = launchBatchSM<decltype(sm)::value>(a); });
    if (e != hipSuccess) return e;
    for (uint32_t k = 0; k < slots; ++k)
        hipLaunchKernelGGL(kBatchGather, dim3((pixels + 255u) / 256u), dim3(256), 0, st, tab,
                           pixPrefix, nActive, k, nullptr, nullptr, 0u);
    return hipGetLastError();
}

hipError_t launchBatchGatherPass(const BatchScene* tab, const uint32_t* pixPrefix,
                                 uint32_t nActive, uint32_t pixels, uint32_t slots,
                                 const float4* L, const float2* P, hipStream_t st) {
    for (uint32_t k = 0; k < slots; ++k)
        hipLaunchKernelGGL(kBatchGather, dim3((pixels + 255u) / 256u), dim3(256), 0, st, tab,
                           pixPrefix, nActive, k, L, P, pixels);
    return hipGetLastError();
}

hipError_t launchFilmGather(const float4* L, const float2* P, uint32_t W,
                            uint32_t H, float* sum, float* wsum,
                            hipStream_t st, uint32_t slots) {
    uint32_t n = W * H;
    if (n == 0 || slots == 0) return hipSuccess;
    hipLaunchKernelGGL(kFilmGather, dim3((n + 255u) / 256u), dim3(256), 0, st,
                       L, P, W, H, slots, sum, wsum);
    return hipGetLastError();
}

}  // namespace crh
