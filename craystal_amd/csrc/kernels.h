// Host-side declarations of the gfx950 kernel launchers (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/crayhip.h"
#include "layout.h"

namespace crh {

constexpr uint32_t kTraceBlock = 256;
constexpr uint32_t kRenderBlock = 256;

hipError_t launchTraceClosest(const SceneView& s, const crh_ray* rays,
                              uint64_t n, crh_hit* hits, uint32_t* work,
                              hipStream_t st);
hipError_t launchTraceAny(const SceneView& s, const crh_ray* rays, uint64_t n,
                          uint8_t* occ, hipStream_t st);
hipError_t launchRenderPass(const SceneView& s, const CameraView& cam,
                            const IntegratorView& cfg, uint32_t si, float4* L,
                            float2* P, unsigned long long* counters,
                            bool countWork, hipStream_t st);
hipError_t launchFilmGather(const float4* L, const float2* P, uint32_t W,
                            uint32_t H, float* sum, float* wsum,
                            hipStream_t st);

}  // namespace crh
