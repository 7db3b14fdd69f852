// Host-side declarations of the gfx950 kernel launchers (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/crayhip.h"
#include "layout.h"
#include "traverse_limits.h"

namespace crh {

constexpr uint32_t kTraceBlock = 256;   // upper bound; launches may use fewer
constexpr uint32_t kRenderBlock = 256;

// Launch choices that do not change results (set by the autotuner).
struct Tuning {
    // ray-batch kernels (crh_trace_*)
    uint32_t privStack = 1;     // 1: traversal stack in private memory
    uint32_t traceBlock = 64;   // threads per block (64..256)
    uint32_t persistent = 0;    // 1: lanes refill with new rays from a work counter
    uint32_t refill = 16;       // idle lanes per wave that trigger a refill
    uint32_t ptGrid = 2048;     // blocks of the persistent variant
    // megakernel
    uint32_t renderPriv = 0;
    // wavefront
    uint32_t wfPriv = 0;
    uint32_t wfTraceGrid = 16384;
    uint32_t wfTraceBlock = 128;
    uint32_t wfShadeGrid = 16384;
    uint32_t wfPersistent = 0;
    uint32_t wfRefill = 16;
};

// counter: one device uint32 used by the persistent variants (may be null).
hipError_t launchTraceClosest(const SceneView& s, const crh_ray* rays,
                              uint64_t n, crh_hit* hits, uint32_t* work,
                              const Tuning& t, uint32_t* counter, hipStream_t st);
hipError_t launchTraceAny(const SceneView& s, const crh_ray* rays, uint64_t n,
                          uint8_t* occ, const Tuning& t, uint32_t* counter, hipStream_t st);
hipError_t launchRenderPass(const SceneView& s, const CameraView& cam,
                            const IntegratorView& cfg, uint32_t si, float4* L,
                            float2* P, unsigned long long* counters,
                            bool countWork, const Tuning& t, hipStream_t st);
hipError_t launchFilmGather(const float4* L, const float2* P, uint32_t W,
                            uint32_t H, float* sum, float* wsum,
                            hipStream_t st);

}  // namespace crh
