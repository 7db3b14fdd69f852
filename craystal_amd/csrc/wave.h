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

// Hands out the indices [0, n) to the lanes of one wave: the wave takes
// chunks of `chunk` consecutive indices from *counter (zero at launch) with
// one atomic per chunk, and gives its lanes that want work the next indices
// of its chunk in lane order (ballot + mbcnt).  All members are
// wave-uniform.  A wave that needs more than its chunk holds takes the rest
// of its chunk and the start of the next one.
// Chunk size for n items over the launch's waves: about four chunks per
// wave, between 64 and 512 items.
CRH_D uint32_t dispenserChunk(uint32_t n) {
    const uint32_t waves = gridDim.x * ((blockDim.x + 63u) / 64u);
    return min(512u, max(64u, n / (4u * waves)));
}

struct WaveDispenser {
    uint32_t next = 0, end = 0;
    bool exhausted = false;  // a chunk past n was drawn: nothing left anywhere

    // Returns the lane's index, or n (or more) when it gets none.
    CRH_D uint32_t take(uint32_t* counter, uint32_t n, uint32_t chunk, bool want) {
        const uint64_t mask = __ballot(want);
        if (mask == 0 || exhausted) return n;
        const uint32_t cnt = uint32_t(__popcll(mask));
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
            uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
        const uint32_t have = end - next;
        uint32_t idx;
        if (have >= cnt) {
            idx = next + rank;
            next += cnt;
        } else {
            const int leader = __ffsll((unsigned long long)mask) - 1;
            uint32_t g = 0;
            if (int(__lane_id()) == leader) g = atomicAdd(counter, chunk);
            g = __shfl(g, leader, 64);
            idx = rank < have ? next + rank : (g >= n ? n : g + (rank - have));
            if (g >= n) {
                exhausted = true;
                next = end = 0;
            } else {
                end = min(g + chunk, n);
                next = min(g + (cnt - have), end);
            }
        }
        return want && idx < n ? idx : n;
    }
};

}  // namespace crh
