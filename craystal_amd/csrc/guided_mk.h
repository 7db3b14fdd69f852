// The fused guided path tracer (guided_mk.hip): persistent path-regenerating
// megakernel whose waves evaluate the guide on MFMA between the light sample
// and the continuation of each bounce.
#pragma once
#include <hip/hip_runtime.h>

#include "guiding.h"
#include "guiding_kernels.h"
#include "kernels.h"

namespace crh {

// Per-wave output buffer of the network for one pass: 64 lanes x kGuideRaw
// floats, for every wave of the persistent grid.
constexpr uint32_t kGuideRaw = 40;
size_t guidedScratchFloats(uint32_t grid);
// Dynamic LDS of one block: the traversal stack, then one activation tile
// per wave (16 rows of width + 8 bf16).
size_t guidedLdsBytes(uint32_t stackMode, uint32_t stackCap, uint32_t width);

struct GuidedPassArgs {
    const GuideView* guide;
    float alpha;
    uint32_t mask;        // bit t: guide BSDF type t (0 Lambertian, 1 Principled)
    uint32_t maxDepth;    // guide vertices at depth < maxDepth (0: all)
    float* scratch;       // guidedScratchFloats(grid) floats
    float4* rec;          // training records (kRecF4 float4 each), may be null
    uint32_t* recCounter; // records written this pass (zeroed by the launcher)
    uint32_t recCap;
};

// Several camera views of one scene in one launch: cams, tilePrefix (16 x 16
// tiles of the views before each, then the total) and pixPrefix (the same
// for pixels) are device arrays; tiles = tilePrefix[n], totalPix =
// pixPrefix[n].
struct GuidedViews {
    const CameraView* cams;
    const uint32_t* tilePrefix;
    const uint32_t* pixPrefix;
    uint32_t n, tiles, totalPix;
};

// Sample indices si .. si + active - 1 into slots of L and P (as
// launchRenderPass) with guided continuations; grid = min(tiles * active,
// t.renderGrid) persistent blocks of 256 threads; next: a device counter.
// views: several views (cam unused), sample slot k of view j's pixel i at
// k * totalPix + pixPrefix[j] + i of L and P.
hipError_t launchGuidedPass(const SceneView& s, const CameraView& cam, const IntegratorView& cfg,
                            uint32_t si, uint32_t active, float4* L, float2* P,
                            unsigned long long* counters, const Tuning& t, uint32_t* next,
                            const GuidedPassArgs& g, hipStream_t st,
                            const GuidedViews* views = nullptr);

// One launch of the fused kernel; launchGuidedSM<SM> (defined once per
// stack mode, guided_mk.hip built with CRH_GMK_PART = SM) picks the register
// budget and MIS variant.
struct GuidedLaunch {
    SceneView s;
    CameraView cam;
    IntegratorView cfg;
    uint32_t si, tilesX, tiles, total;
    uint32_t* next;
    float4* L;
    float2* P;
    unsigned long long* counters;
    const GuidedPassArgs* g;
    uint32_t grid, tileOff;
    size_t lds;
    uint32_t waves;
    hipStream_t st;
    GuidedViews views;
    bool multi;
};
template <int SM> hipError_t launchGuidedSM(const GuidedLaunch& a);
template <> hipError_t launchGuidedSM<0>(const GuidedLaunch&);
template <> hipError_t launchGuidedSM<1>(const GuidedLaunch&);
template <> hipError_t launchGuidedSM<2>(const GuidedLaunch&);

// Training samples from the pass's records (as the wavefront's): incident
// radiance along the guided direction from the final radiance L[out] of the
// record's path.  trainCounter receives the samples produced.
hipError_t launchGuidedRecordsToSamples(const float4* rec, const uint32_t* recCounter,
                                        uint32_t recCap, const float4* L, GuideSample* out,
                                        uint32_t cap, uint32_t* trainCounter,
                                        uint32_t radianceTarget, hipStream_t st);

}  // namespace crh
