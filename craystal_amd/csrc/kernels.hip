// gfx950 kernels: batch closest-hit / any-hit queries, the plain megakernel
// integrator (one path per thread per sample index) and the deterministic
// film gather.  Host launchers at the bottom are called by api.cpp.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "integrator.h"
#include "path.h"
#include "wave.h"

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

// ---------------------------------------------------------------------------
// Megakernel integrator (Core/Integrator.cu:142-227); per-path steps in path.h
// ---------------------------------------------------------------------------
// Workgroups go to the 8 XCDs round robin (block b to XCD b mod 8), and
// each XCD has its own L2.  This bijection on [0, n) gives the blocks of
// one XCD a contiguous run of work items (a band of tiles of the image), so
// neighbouring primary rays share their XCD's L2; the last n mod 8 items
// map to themselves.
CRH_D uint32_t xcdBand(uint32_t b, uint32_t n) {
    const uint32_t per = n / 8u;
    if (b >= 8u * per) return b;
    return (b % 8u) * per + b / 8u;
}

CRH_D void tileCoord(uint32_t tile, uint32_t tilesX, uint32_t& x, uint32_t& y) {
    tileCoordOf(tile, threadIdx.x, tilesX, x, y);
}

// pathTraceKernel (Core/Integrator.cu:142-227) for one pixel: radiance of
// sample index `si` through pixel (x, y) into L[id] (float4, w unused) and
// the jittered film position into P[id]; adds the rays traced to `rays`.
// SPEC: spectral transport; the path draws its hero wavelength bin right
// after the pixel jitter, and L receives CIE XYZ instead of RGB.
template <bool MIS, bool COUNT, bool SPEC, class Stack>
CRH_D void tracePixel(const SceneView& s, Stack& stk, const CameraView& cam,
                      const IntegratorView& cfg, uint32_t x, uint32_t y, uint32_t si,
                      float4* __restrict__ L, float2* __restrict__ P, uint32_t& rays,
                      WorkCount& wc) {
    PathCtx<COUNT, Stack> c{s, stk, 0, {0, 0}};
    Sampler smp(x, y, si);
    F2 j = smp.next2D();
    F2 pix{float(x) + j.x, float(y) + j.y};
    const Lanes w = SPEC ? heroLanes(s.spec, smp.next1D()) : rgbLanes();
    Ray ray = cameraRay(cam, pix);
    Spec rad = sp(0.f), beta = sp(1.f);
    for (uint32_t depth = 0; depth < cfg.maxDepth; ++depth)
        if (!pathBounce<MIS, SPEC>(c, cfg, depth, ray, rad, beta, smp, w)) break;
    if (SPEC) rad = laneXYZ(rad, w);
    size_t id = size_t(y) * cam.width + x;
    L[id] = make_float4(rad.r, rad.g, rad.b, 0.f);
    P[id] = make_float2(pix.x, pix.y);
    rays += c.rays;
    wc.nodes += c.wc.nodes;
    wc.tris += c.wc.tris;
}

// Wave-level reductions, one atomic per wave and counter:
// counters = {rays, node visits, triangle tests}.
template <bool COUNT>
CRH_D void flushCounters(unsigned long long* counters, uint32_t rays, uint32_t nodes,
                         uint32_t tris) {
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

// Work item w = xcdBand(block) renders tile w % tiles of sample index
// si + w / tiles into slot w / tiles of L and P (npix entries per slot).
// W: minimum waves per SIMD the register allocation must allow (1 = the
// compiler's choice; 4 caps the kernel at 128 VGPRs).
template <bool MIS, bool COUNT, int SM, int W, bool SPEC = false>
__global__ __launch_bounds__(kRenderBlock) __attribute__((amdgpu_waves_per_eu(W, 8))) void kRenderPass(
    SceneView s, CameraView cam, IntegratorView cfg, uint32_t si,
    uint32_t tilesX, uint32_t tiles, uint32_t band, float4* __restrict__ L,
    float2* __restrict__ P, unsigned long long* __restrict__ counters) {
    extern __shared__ uint2 lds[];
    const uint32_t b = band ? xcdBand(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t slot = b / tiles;
    const size_t off = size_t(slot) * cam.width * cam.height;
    uint32_t x, y;
    tileCoord(b - slot * tiles, tilesX, x, y);
    uint32_t rays = 0;
    WorkCount wc{0, 0};
    if (x < cam.width && y < cam.height) {
        auto stk = makeStack<SM>(lds);
        tracePixel<MIS, COUNT, SPEC>(s, stk, cam, cfg, x, y, si + slot, L + off, P + off, rays,
                                     wc);
    }
    flushCounters<COUNT>(counters, rays, wc.nodes, wc.tris);
}

// Path regeneration: persistent blocks; every lane runs one bounce of its
// path per iteration and, when its path ends, writes it and takes the next
// path index (a WaveDispenser: one atomic per chunk of paths), so lanes do
// not idle while the longest path of their wave finishes.  Path p is pixel
// p % (tiles * 256) (tile order, 8x8 per wave as kRenderPass) of sample
// slot p / (tiles * 256).  Each path runs exactly tracePixel's steps on the
// same random stream, so L and P equal kRenderPass's.
template <bool MIS, bool SPEC, int SM, int W>
__global__ __launch_bounds__(kRenderBlock) __attribute__((amdgpu_waves_per_eu(W, 8))) void
kRenderRegen(SceneView s, CameraView cam, IntegratorView cfg, uint32_t si, uint32_t tilesX,
             uint32_t tiles, uint32_t total, uint32_t* __restrict__ next,
             float4* __restrict__ L, float2* __restrict__ P,
             unsigned long long* __restrict__ counters) {
    extern __shared__ uint2 lds[];
    auto stk = makeStack<SM>(lds);
    PathCtx<false, typename StackOf<SM>::T> c{s, stk, 0, {0, 0}};
    const size_t npix = size_t(cam.width) * cam.height;
    const uint32_t perSlot = tiles * 256u;
    bool live = false;
    WaveDispenser disp;
    const uint32_t chunk = dispenserChunk(total);
    uint32_t depth = 0;
    size_t out = 0;
    Sampler smp(0u, 0u, 0u);
    Ray ray{};
    Spec rad = sp(0.f), beta = sp(1.f);
    F2 pix{0.f, 0.f};
    Lanes w = rgbLanes();
    auto finish = [&]() {
        Spec v = SPEC ? laneXYZ(rad, w) : rad;
        L[out] = make_float4(v.r, v.g, v.b, 0.f);
        P[out] = make_float2(pix.x, pix.y);
        live = false;
    };
    while (true) {
        {
            const uint32_t p = disp.take(next, total, chunk, !live);
            if (p < total) {
                const uint32_t slot = p / perSlot, q = p - slot * perSlot;
                uint32_t x, y;
                tileCoordOf(q >> 8, q & 255u, tilesX, x, y);
                if (x < cam.width && y < cam.height) {
                    smp = Sampler(x, y, si + slot);
                    F2 j = smp.next2D();
                    pix = F2{float(x) + j.x, float(y) + j.y};
                    w = SPEC ? heroLanes(s.spec, smp.next1D()) : rgbLanes();
                    ray = cameraRay(cam, pix);
                    rad = sp(0.f);
                    beta = sp(1.f);
                    depth = 0;
                    out = slot * npix + size_t(y) * cam.width + x;
                    live = true;
                    if (cfg.maxDepth == 0u) finish();
                }
            }
        }
        if (!__any(live)) {
            if (disp.exhausted) break;
            continue;
        }
        if (live) {
            bool done = !pathBounce<MIS, SPEC>(c, cfg, depth, ray, rad, beta, smp, w);
            if (!done && ++depth >= cfg.maxDepth) done = true;
            if (done) finish();
        }
    }
    flushCounters<false>(counters, c.rays, 0, 0);
}

__global__ void kResetNext(uint32_t* __restrict__ next) {
    if (threadIdx.x == 0) *next = 0u;
}

// Deterministic film update for one pass.  Each pixel gathers the tent-filter
// contributions of the 3x3 neighbouring samples in row-major sample order,
// which is the order a sequential SensorData::addSample loop over the pass
// (Core/Sensor.cu:19-43) delivers them; samples with NaN/Inf are dropped and
// zero weights skipped (they add exactly nothing).
CRH_D void gatherPixel(const float4* __restrict__ L, const float2* __restrict__ P,
                       uint32_t W, uint32_t H, float* __restrict__ sum,
                       float* __restrict__ wsum, uint32_t i) {
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

__global__ __launch_bounds__(256) void kFilmGather(
    const float4* __restrict__ L, const float2* __restrict__ P, uint32_t W,
    uint32_t H, float* __restrict__ sum, float* __restrict__ wsum) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= W * H) return;
    gatherPixel(L, P, W, H, sum, wsum, i);
}

// ---------------------------------------------------------------------------
// Batched rendering: many scenes, one persistent launch per sample index
// ---------------------------------------------------------------------------
// The table holds the active scenes first (sorted by decreasing spp, so the
// scenes that still render at pass k are a prefix of it).  tilePrefix[j] is
// the first work item (tile) of scene j; tilePrefix[nActive] = item count.

// Index j with prefix[j] <= item < prefix[j + 1] (prefix[0] = 0).
CRH_D uint32_t findScene(const uint32_t* __restrict__ prefix, uint32_t n, uint32_t item) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= item) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void kBatchReset(uint32_t* __restrict__ next) {
    if (threadIdx.x < kBatchMaxChunks) next[threadIdx.x] = 0u;
}

// Persistent blocks take tiles from `chunks` contiguous ranges of the item
// list, each with its own counter; block b starts on range b % chunks (the
// hardware deals blocks to the 8 XCDs round-robin, so with 8 ranges each
// XCD's L2 sees mostly its own scenes) and then moves on to the others.
// Every block leaves after one failed take per range.
template <bool MIS, int SM, int W>
__global__ __launch_bounds__(kRenderBlock) __attribute__((amdgpu_waves_per_eu(W, 8))) void kBatchRender(
    const BatchScene* __restrict__ tab, const uint32_t* __restrict__ tilePrefix,
    uint32_t nActive, IntegratorView cfg, uint32_t si, uint32_t slots, uint32_t tileW,
    uint32_t chunks, uint32_t* __restrict__ next, unsigned long long* __restrict__ counters) {
    extern __shared__ uint2 lds[];
    __shared__ uint32_t taken;
    // items: slot-major, slot k's tiles of every active scene, then slot k + 1's
    const uint32_t tilesPerSlot = tilePrefix[nActive];
    const uint32_t T = tilesPerSlot * slots;
    const uint32_t tileH = kRenderBlock / tileW;
    const uint32_t home = blockIdx.x % chunks;
    uint32_t rays = 0;
    WorkCount wc{0, 0};
    auto stk = makeStack<SM>(lds);
    for (uint32_t k = 0; k < chunks; ++k) {
        const uint32_t c = (home + k) % chunks;
        const uint32_t lo = uint32_t(uint64_t(T) * c / chunks);
        const uint32_t hi = uint32_t(uint64_t(T) * (c + 1) / chunks);
        for (;;) {
            if (threadIdx.x == 0) taken = lo + atomicAdd(&next[c], 1u);
            __syncthreads();
            const uint32_t item = __builtin_amdgcn_readfirstlane(taken);
            __syncthreads();
            if (item >= hi) break;
            const uint32_t slot = item / tilesPerSlot;
            const uint32_t rem = item - slot * tilesPerSlot;
            const uint32_t j = __builtin_amdgcn_readfirstlane(findScene(tilePrefix, nActive, rem));
            const BatchScene& b = tab[j];
            const uint32_t tile = rem - tilePrefix[j];
            const uint32_t x = (tile % b.tilesX) * tileW + threadIdx.x % tileW;
            const uint32_t y = (tile / b.tilesX) * tileH + threadIdx.x / tileW;
            const size_t off = size_t(slot) * b.cam.width * b.cam.height;
            if (x < b.cam.width && y < b.cam.height)
                tracePixel<MIS, false, false>(b.view, stk, b.cam, cfg, x, y, si + slot,
                                       reinterpret_cast<float4*>(b.L) + off,
                                       reinterpret_cast<float2*>(b.P) + off, rays, wc);
        }
    }
    flushCounters<false>(counters, rays, 0, 0);
}

// Film update of every active scene with the samples of slot `slot`:
// pixPrefix[j] is scene j's first pixel.
__global__ __launch_bounds__(256) void kBatchGather(const BatchScene* __restrict__ tab,
                                                    const uint32_t* __restrict__ pixPrefix,
                                                    uint32_t nActive, uint32_t slot) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= pixPrefix[nActive]) return;
    const uint32_t j = findScene(pixPrefix, nActive, i);
    const BatchScene& b = tab[j];
    const size_t off = size_t(slot) * b.cam.width * b.cam.height;
    gatherPixel(reinterpret_cast<const float4*>(b.L) + off,
                reinterpret_cast<const float2*>(b.P) + off, b.cam.width, b.cam.height, b.sum,
                b.wsum, i - pixPrefix[j]);
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
    const size_t lds = stackLdsBytes(t.renderPriv, s.stackCap, kRenderBlock);
    const dim3 block(kRenderBlock);
    if (t.renderRegen && !countWork && next) {
        const uint32_t total = tiles * active * 256u;
        const dim3 grid(std::min<uint32_t>(tiles * active, t.renderGrid));
        hipLaunchKernelGGL(kResetNext, dim3(1), dim3(64), 0, st, next);
        withStack(t.renderPriv, [&](auto sm) {
            constexpr int SM = decltype(sm)::value;
            withWaves(t.renderWaves, [&](auto wv) {
                constexpr int W = decltype(wv)::value;
                if (spectral && cfg.useMIS)
                    hipLaunchKernelGGL((kRenderRegen<true, true, SM, W>), grid, block, lds, st, s,
                                       cam, cfg, si, tilesX, tiles, total, next, L, P, counters);
                else if (spectral)
                    hipLaunchKernelGGL((kRenderRegen<false, true, SM, W>), grid, block, lds, st,
                                       s, cam, cfg, si, tilesX, tiles, total, next, L, P,
                                       counters);
                else if (cfg.useMIS)
                    hipLaunchKernelGGL((kRenderRegen<true, false, SM, W>), grid, block, lds, st,
                                       s, cam, cfg, si, tilesX, tiles, total, next, L, P,
                                       counters);
                else
                    hipLaunchKernelGGL((kRenderRegen<false, false, SM, W>), grid, block, lds, st,
                                       s, cam, cfg, si, tilesX, tiles, total, next, L, P,
                                       counters);
            });
        });
        return hipGetLastError();
    }
    const dim3 grid(tiles * active);
    const uint32_t band = t.xcdBand;
    withStack(t.renderPriv, [&](auto sm) {
        constexpr int SM = decltype(sm)::value;
        withWaves(t.renderWaves, [&](auto wv) {
            constexpr int W = decltype(wv)::value;
            if (spectral && cfg.useMIS)
                hipLaunchKernelGGL((kRenderPass<true, false, SM, W, true>), grid, block, lds, st, s,
                                   cam, cfg, si, tilesX, tiles, band, L, P, counters);
            else if (spectral)
                hipLaunchKernelGGL((kRenderPass<false, false, SM, W, true>), grid, block, lds, st,
                                   s, cam, cfg, si, tilesX, tiles, band, L, P, counters);
            else if (cfg.useMIS && countWork)
                hipLaunchKernelGGL((kRenderPass<true, true, SM, W>), grid, block, lds, st, s, cam,
                                   cfg, si, tilesX, tiles, band, L, P, counters);
            else if (cfg.useMIS)
                hipLaunchKernelGGL((kRenderPass<true, false, SM, W>), grid, block, lds, st, s, cam,
                                   cfg, si, tilesX, tiles, band, L, P, counters);
            else if (countWork)
                hipLaunchKernelGGL((kRenderPass<false, true, SM, W>), grid, block, lds, st, s, cam,
                                   cfg, si, tilesX, tiles, band, L, P, counters);
            else
                hipLaunchKernelGGL((kRenderPass<false, false, SM, W>), grid, block, lds, st, s,
                                   cam, cfg, si, tilesX, tiles, band, L, P, counters);
        });
    });
    return hipGetLastError();
}

hipError_t launchBatchPass(const BatchScene* tab, const uint32_t* tilePrefix,
                           const uint32_t* pixPrefix, uint32_t nActive, uint32_t items,
                           uint32_t pixels, const IntegratorView& cfg, uint32_t si,
                           uint32_t slots, uint32_t stackCap, const Tuning& t, uint32_t* next,
                           unsigned long long* counters, hipStream_t st) {
    if (nActive == 0 || items == 0 || slots == 0) return hipSuccess;
    if (uint64_t(items) * slots >= (uint64_t(1) << 32)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kBatchReset, dim3(1), dim3(64), 0, st, next);
    const size_t lds = stackLdsBytes(t.batchPriv, stackCap, kRenderBlock);
    const uint32_t total = items * slots;
    const dim3 grid(total < t.batchGrid ? total : t.batchGrid), block(kRenderBlock);
    const uint32_t tw = t.batchTileW, ch = t.batchChunks;
    withStack(t.batchPriv, [&](auto sm) {
        constexpr int SM = decltype(sm)::value;
        withWaves(t.batchWaves, [&](auto wv) {
            constexpr int W = decltype(wv)::value;
            if (cfg.useMIS)
                hipLaunchKernelGGL((kBatchRender<true, SM, W>), grid, block, lds, st, tab,
                                   tilePrefix, nActive, cfg, si, slots, tw, ch, next, counters);
            else
                hipLaunchKernelGGL((kBatchRender<false, SM, W>), grid, block, lds, st, tab,
                                   tilePrefix, nActive, cfg, si, slots, tw, ch, next, counters);
        });
    });
    for (uint32_t k = 0; k < slots; ++k)
        hipLaunchKernelGGL(kBatchGather, dim3((pixels + 255u) / 256u), dim3(256), 0, st, tab,
                           pixPrefix, nActive, k);
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
