// Device helper shared by the batched kernels (the batched megakernel, its
// film gather and the batched wavefront): which scene of a batch an item
// belongs to.
#pragma once
#include "device_math.h"

namespace crh {

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

}  // namespace crh
