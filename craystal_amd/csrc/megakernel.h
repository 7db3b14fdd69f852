// Device code of the megakernels (the plain path tracer, its path-regenerating
// variant and the batched many-scene launch), shared by the per-stack-mode
// translation units mk_inst.hip is compiled into.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "integrator.h"
#include "path.h"
#include "wave.h"
#include "batch_device.h"

namespace crh {

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


// ---------------------------------------------------------------------------
// Batched rendering: many scenes, one persistent launch per sample index
// ---------------------------------------------------------------------------
// The table holds the active scenes first (sorted by decreasing spp, so the
// scenes that still render at pass k are a prefix of it).  tilePrefix[j] is
// the first work item (tile) of scene j; tilePrefix[nActive] = item count.

// Persistent blocks take tiles from `chunks` contiguous ranges of the item
// list, each with its own counter; block b starts on range b % chunks (the
// hardware deals blocks to the 8 XCDs round-robin, so with 8 ranges each
// XCD's L2 sees mostly its own scenes) and then moves on to the others.
// Every block leaves after one failed take per range.
template <bool MIS, bool SPEC, int SM, int W>
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
                tracePixel<MIS, false, SPEC>(b.view, stk, b.cam, cfg, x, y, si + slot,
                                       reinterpret_cast<float4*>(b.L) + off,
                                       reinterpret_cast<float2*>(b.P) + off, rays, wc);
        }
    }
    flushCounters<false>(counters, rays, 0, 0);
}


}  // namespace crh
