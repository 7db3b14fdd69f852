// Host-side declarations of the gfx950 kernel launchers (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/crayhip.h"
#include "layout.h"
#include "traverse_limits.h"

#include <type_traits>

namespace crh {

constexpr uint32_t kTraceBlock = 256;   // upper bound; launches may use fewer
constexpr uint32_t kRenderBlock = 256;

// Launch choices that do not change results (set by the autotuner).
struct Tuning {
    // Traversal stack mode of each kernel family: 0 LDS, 1 private, 2 short
    // LDS stack with private overflow.
    // ray-batch kernels (crh_trace_*)
    uint32_t privStack = 2;
    uint32_t traceBlock = 64;   // threads per block (64..256)
    uint32_t persistent = 0;    // 1: lanes refill with new rays from a work counter
    uint32_t refill = 16;       // idle lanes per wave that trigger a refill
    uint32_t ptGrid = 2048;     // blocks of the persistent variant
    // megakernel
    uint32_t renderPriv = 2;
    // wavefront
    uint32_t wfPriv = 2;
    uint32_t wfTraceGrid = 16384;
    uint32_t wfTraceBlock = 128;
    uint32_t wfShadeGrid = 16384;
    uint32_t wfPersistent = 0;
    uint32_t wfRefill = 16;
    uint32_t wfPool = 1u << 23;  // paths in flight (several samples per pixel for small images)
    uint32_t renderPool = 1u << 23;  // megakernel: paths per launch, the same way
    uint32_t renderWaves = 8;        // megakernels: minimum waves per SIMD (1, 4, 5, 6 or 8)
    uint32_t batchWaves = 8;         // batch megakernel: the same for kBatchRender
    uint32_t batchSlots = 16;        // batch: sample indices per launch
    uint32_t traceTris = 1;          // persistent ray batches: triangles per step
    uint32_t renderRegen = 1;        // megakernel: lanes take a new path when theirs ends
    uint32_t renderGrid = 2048;      // megakernel with regeneration: persistent blocks
    uint32_t xcdBand = 0;            // megakernel: each XCD takes a contiguous band of tiles
    uint32_t wfFinish = 2;           // wavefront: finish per path after this many bounces (0 off)
    // batched rendering
    uint32_t batchPriv = 2;
    uint32_t batchTileW = 16;   // tile width in pixels; height = 256 / width
    uint32_t batchGrid = 1024;  // persistent blocks
    uint32_t batchChunks = 8;   // item ranges with their own counters (1..8)
    uint32_t batchGraph = 1;    // 1: passes replayed from a captured hipGraph
    // fused guided megakernel
    uint32_t guidedFused = 1;         // 1: guided rendering in the fused megakernel
    uint32_t guidedPriv = 2;          // stack mode
    uint32_t guidedWaves = 6;         // register budget (waves per SIMD)
    uint32_t guidedGrid = 4096;       // persistent blocks
    uint32_t guidedPool = 1u << 23;   // paths per launch
};

constexpr uint32_t kBatchMaxChunks = 8;

// Traversal stack modes of the tuning: 0 LDS, 1 private, 2 short LDS stack
// with private overflow.  withStack calls f(std::integral_constant<int, mode>).
template <class F>
inline void withStack(uint32_t mode, F&& f) {
    if (mode == 1) f(std::integral_constant<int, 1>{});
    else if (mode == 2) f(std::integral_constant<int, 2>{});
    else f(std::integral_constant<int, 0>{});
}
// Register budget of the megakernels: 1 = the compiler's choice, 4 = at most
// 128 VGPRs (4 waves per SIMD), 5 = at most 96 VGPRs (5 waves, more spills),
// 6 = at most 80 VGPRs, 8 = at most 64 (the most waves a SIMD holds).
// withWaves calls f(integral_constant<int, w>).
template <class F>
inline void withWaves(uint32_t w, F&& f) {
    if (w == 4) f(std::integral_constant<int, 4>{});
    else if (w == 5) f(std::integral_constant<int, 5>{});
    else if (w == 6) f(std::integral_constant<int, 6>{});
    else if (w == 8) f(std::integral_constant<int, 8>{});
    else f(std::integral_constant<int, 1>{});
}
// LDS bytes per block of a stack mode.
size_t stackLdsBytes(uint32_t mode, uint32_t stackCap, uint32_t block);

// counter: one device uint32 used by the persistent variants (may be null).
hipError_t launchTraceClosest(const SceneView& s, const crh_ray* rays,
                              uint64_t n, crh_hit* hits, uint32_t* work,
                              const Tuning& t, uint32_t* counter, hipStream_t st);
hipError_t launchTraceAny(const SceneView& s, const crh_ray* rays, uint64_t n,
                          uint8_t* occ, const Tuning& t, uint32_t* counter, hipStream_t st);
// Sample indices si .. si + active - 1 in one launch; slot k of L and P
// (w * h entries each) receives sample si + k.  next: a device counter for
// path regeneration (used when the tuning asks for it and work is not
// counted).
hipError_t launchRenderPass(const SceneView& s, const CameraView& cam,
                            const IntegratorView& cfg, uint32_t si, uint32_t active, float4* L,
                            float2* P, unsigned long long* counters,
                            bool countWork, const Tuning& t, hipStream_t st,
                            bool spectral = false, uint32_t* next = nullptr);
// Sample indices si .. si + slots - 1 over the first nActive scenes of the
// table: one persistent megakernel launch over `items` tiles of each of the
// `slots` sample indices (slot k of a scene's film buffers receives sample
// si + k), then one film gather per sample index, in sample order, over
// `pixels` pixels.  items * slots must fit in 32 bits.  next:
// kBatchMaxChunks device counters.
hipError_t launchBatchPass(const BatchScene* tab, const uint32_t* tilePrefix,
                           const uint32_t* pixPrefix, uint32_t nActive, uint32_t items,
                           uint32_t pixels, const IntegratorView& cfg, uint32_t si,
                           uint32_t slots, uint32_t stackCap, const Tuning& t, uint32_t* next,
                           unsigned long long* counters, hipStream_t st, bool spectral = false);
// Arguments of one megakernel launch (kRenderPass or kRenderRegen) and the
// per-stack-mode launchers: mk_inst.hip is compiled once per kernel family
// and stack mode, each defining one of the specialisations below.
struct MkArgs {
    SceneView s;
    CameraView cam;
    IntegratorView cfg;
    uint32_t si, tilesX, tiles, band, total;  // total: paths of a regenerating launch
    uint32_t* next;
    float4* L;
    float2* P;
    unsigned long long* counters;
    dim3 grid;
    size_t lds;
    uint32_t waves;
    bool mis, count, spectral;
    hipStream_t st;
};
struct BatchArgs {
    const BatchScene* tab;
    const uint32_t* tilePrefix;
    uint32_t nActive;
    IntegratorView cfg;
    uint32_t si, slots, tileW, chunks;
    uint32_t* next;
    unsigned long long* counters;
    dim3 grid;
    size_t lds;
    uint32_t waves;
    bool mis, spectral;
    hipStream_t st;
};
template <int SM> hipError_t launchRenderSM(const MkArgs& a);
template <int SM> hipError_t launchRegenSM(const MkArgs& a);
template <int SM> hipError_t launchBatchSM(const BatchArgs& a);
template <> hipError_t launchRenderSM<0>(const MkArgs&);
template <> hipError_t launchRenderSM<1>(const MkArgs&);
template <> hipError_t launchRenderSM<2>(const MkArgs&);
template <> hipError_t launchRegenSM<0>(const MkArgs&);
template <> hipError_t launchRegenSM<1>(const MkArgs&);
template <> hipError_t launchRegenSM<2>(const MkArgs&);
template <> hipError_t launchBatchSM<0>(const BatchArgs&);
template <> hipError_t launchBatchSM<1>(const BatchArgs&);
template <> hipError_t launchBatchSM<2>(const BatchArgs&);

// Film updates of the first nActive scenes of the table from a pass buffer
// holding all of them (the batched wavefront's): slot k of scene j at
// L + k * pixels + pixPrefix[j], pixels = pixPrefix[nActive]; one launch per
// slot, in sample order.
hipError_t launchBatchGatherPass(const BatchScene* tab, const uint32_t* pixPrefix,
                                 uint32_t nActive, uint32_t pixels, uint32_t slots,
                                 const float4* L, const float2* P, hipStream_t st);
// Film update with `slots` consecutive sample slots (slot k at L + k * W * H),
// in slot order.
hipError_t launchFilmGather(const float4* L, const float2* P, uint32_t W,
                            uint32_t H, float* sum, float* wsum,
                            hipStream_t st, uint32_t slots = 1);

}  // namespace crh
