// Wave-level (64-lane) helpers shared by the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "device_math.h"

namespace crh {

// Reserve `pred ? 1 : 0` consecutive slots of `counter` for each active lane
// of the wave with one atomic: returns the lane's slot, or 0xFFFFFFFF when
// pred is false.
CRH_D uint32_t wavePush(uint32_t* counter, bool pred) {
    uint64_t mask = __ballot(pred);
    if (mask == 0) return 0xFFFFFFFFu;
    uint32_t lo = uint32_t(mask), hi = uint32_t(mask >> 32);
    uint32_t below = __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
    int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (int(__lane_id()) == leader) base = atomicAdd(counter, uint32_t(__popcll(mask)));
    base = __shfl(base, leader, 64);
    return pred ? base + below : 0xFFFFFFFFu;
}

}  // namespace crh
