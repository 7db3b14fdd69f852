// The allocation-free closest-hit query (noalloc.hip): its scratch lives in a
// file-scope device array of fixed size instead of library allocations.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace crh {

constexpr size_t kNoAllocHeadBytes = 256;          // counter + work totals
constexpr uint32_t kNoAllocChunkRays = 1u << 20;   // per-ray counts held at once (8 MiB)

// Finds this device's copy of the scratch array (hipGetSymbolAddress) and
// keeps its address; call once per device with that device current.
hipError_t noallocScratchInit(int device);
// Bytes of the scratch array (per GPU).
size_t noallocScratchBytes();
// Closest hits of rays[0, n) into hits; with countWork, also the total node
// visits and triangle tests (totals[0], totals[1]; the call then waits for
// the stream).  Shares the kernels of launchTraceClosest.
hipError_t traceClosestNoAlloc(int device, const SceneView& s, const crh_ray* rays, uint64_t n,
                               crh_hit* hits, bool countWork, const Tuning& t, hipStream_t st,
                               unsigned long long totals[2]);

}  // namespace crh
