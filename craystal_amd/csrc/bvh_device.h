// The binned-SAH BVH builder on the GPU (bvh_device.hip): the same BvhResult
// as buildBvh, bit for bit, for the same parameters.
#pragma once
#include <hip/hip_runtime.h>

#include "bvh_build.h"

namespace crh {

constexpr uint32_t kDeviceMaxBins = 64;   // bins per axis the GPU builder takes
constexpr uint32_t kDeviceMaxDepth = 64;  // largest maxDepth it takes

// dPos / dIdx: device arrays (positions xyz, global vertex indices of ntri
// triangles).  Waits for the stream; out.buildMs is the wall time of the call.
hipError_t buildBvhDevice(const float* dPos, const uint32_t* dIdx, uint32_t ntri,
                          const BuildParams& p, hipStream_t st, BvhResult& out);

}  // namespace crh
