// Binned-SAH BVH builder on the GPU (gfx950), bit-exact with buildBvh
// (bvh_build.cpp) and so with BVH::build (Core/BVH.cu:88-338).
//
// Work split.  Tasks (node ranges of the index array) larger than
// kSmallMax triangles are built level by level: the host holds the level's
// task list and launches, over chunks of kChunk triangles, the bounds, the
// binning, one SAH scan per task and the partition; it reads back only the
// left counts.  Every task of at most kSmallMax triangles is then built with
// its whole subtree by one workgroup, depth first, its slice of the index
// array held in LDS.  Nodes are numbered afterwards in the reference's
// post-order by a host walk over both node sets.
//
// Exactness.  The sequential build grows boxes with first-wins compares
// (b < a ? b : a): among equal values the earliest triangle of the range
// wins, which is visible only between -0 and +0.  Here every min / max is
// an atomic over a 64-bit key (order-preserving bits of the value, -0 folded
// onto +0, then the triangle's position in the range, smallest first), and
// the value is read back from that triangle, so the result is the
// sequential one whatever the order of the atomics.  The SAH scan, split
// position, bin index (with the x86 float -> int conversion) and costs are
// the host's operations in the host's order.  The partition is libstdc++'s
// two-pointer std::partition: it swaps the r-th element that fails the
// predicate counted from the left with the r-th element that passes it
// counted from the right, for r below the number of failing elements left
// of the final split; each element finds its r from a prefix count, and the
// pairs are swapped.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>
#include <vector>

#include "bvh_build.h"
#include "bvh_device.h"

namespace crh {

namespace {

constexpr uint32_t kSmallMax = 4096;  // tasks at most this large: one workgroup per subtree
constexpr uint32_t kChunk = 16384;    // triangles per workgroup of a large task
constexpr uint32_t kStackMax = 72;    // small-task stack entries (maxDepth <= 64)
constexpr uint32_t kDevBit = 0x80000000u;
constexpr unsigned long long kEmptyMin = ~0ull, kEmptyMax = 0ull;

struct TriSoA {
    float* c[3];   // centroids
    float* mn[3];  // triangle bounds
    float* mx[3];
};

struct BigTask {
    uint32_t start, count, depth, pad;
};
struct Chunk {
    uint32_t task, begin, end, pad;  // begin/end: absolute positions in the index array
};
struct SmallTask {
    uint32_t start, count, depth, pad;
};
struct DNode {
    float mn[3], mx[3];
    uint32_t leaf, a, b, pad;
};

__device__ __forceinline__ float fminS(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float fmaxS(float a, float b) { return (a < b) ? b : a; }
// x86-64 cvttss2si: NaN and out-of-range values give INT_MIN.
__device__ __forceinline__ int truncX86(float x) {
    return (x >= -2147483648.0f && x < 2147483648.0f) ? int(x) : INT_MIN;
}
__device__ __forceinline__ uint32_t orderedKey(float v) {
    uint32_t b = __float_as_uint(v);
    if (b == 0x80000000u) b = 0u;  // -0 equals +0: the position decides
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ unsigned long long packMin(float v, uint32_t pos) {
    return (static_cast<unsigned long long>(orderedKey(v)) << 32) | pos;
}
__device__ __forceinline__ unsigned long long packMax(float v, uint32_t pos) {
    return (static_cast<unsigned long long>(orderedKey(v)) << 32) | (0xFFFFFFFFu - pos);
}
__device__ __forceinline__ unsigned long long umin64(unsigned long long a, unsigned long long b) {
    return b < a ? b : a;
}
__device__ __forceinline__ unsigned long long umax64(unsigned long long a, unsigned long long b) {
    return a < b ? b : a;
}
// Value of a min (j < 3) or max (j >= 3) key of coordinate j % 3: the bound
// of the triangle at that position of the range (empty: +inf / -inf).
__device__ __forceinline__ float decodeKey(unsigned long long key, int j, const TriSoA& t,
                                           const uint32_t* ind, uint32_t start) {
    const bool isMax = j >= 3;
    if (key == (isMax ? kEmptyMax : kEmptyMin)) return isMax ? -INFINITY : INFINITY;
    const uint32_t low = uint32_t(key);
    const uint32_t pos = isMax ? 0xFFFFFFFFu - low : low;
    const uint32_t tri = ind[start + pos];
    return isMax ? t.mx[j - 3][tri] : t.mn[j][tri];
}
__device__ __forceinline__ float boxArea(const float* mn, const float* mx) {
    float d0 = mx[0] - mn[0], d1 = mx[1] - mn[1], d2 = mx[2] - mn[2];
    return float(double(d0 * d1 + d0 * d2 + d1 * d2) * 2.0);
}

__device__ __forceinline__ unsigned long long waveMin64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v = umin64(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ unsigned long long waveMax64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v = umax64(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ uint32_t waveSum(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// Exclusive prefix of pred over the 256 lanes of the block (wave ballots,
// then wave totals through LDS); *total receives the block's count.
__device__ __forceinline__ uint32_t blockPrefix(bool pred, uint32_t* sWave, uint32_t* total) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const unsigned long long m = __ballot(pred);
    const uint32_t below = __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32),
                                                     __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
    if (lane == 0) sWave[wave] = uint32_t(__popcll(m));
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (uint32_t w = 0; w < 4; ++w) {
        if (w < wave) off += sWave[w];
        tot += sWave[w];
    }
    __syncthreads();
    *total = tot;
    return off + below;
}

// ---- triangle data -----------------------------------------------------------
// gatherTriangles (Core/BVH.cu:120-148): centroid (v0 + v1 + v2) / 3 and the
// first-wins bounds starting from +-inf.
__global__ __launch_bounds__(256) void kPrep(const float* __restrict__ pos,
                                             const uint32_t* __restrict__ idx, uint32_t n,
                                             TriSoA t, uint32_t* __restrict__ ind) {
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        const float* v0 = pos + 3 * size_t(idx[3 * size_t(i)]);
        const float* v1 = pos + 3 * size_t(idx[3 * size_t(i) + 1]);
        const float* v2 = pos + 3 * size_t(idx[3 * size_t(i) + 2]);
        for (int k = 0; k < 3; ++k) {
            const float a = v0[k], b = v1[k], c = v2[k];
            t.c[k][i] = (a + b + c) / 3.0f;
            t.mn[k][i] = fminS(fminS(fminS(INFINITY, a), b), c);
            t.mx[k][i] = fmaxS(fmaxS(fmaxS(-INFINITY, a), b), c);
        }
        ind[i] = i;
    }
}

// ---- large tasks ---------------------------------------------------------------
// Node bounds: keys of the chunk reduced per wave, then global atomics.
__global__ __launch_bounds__(256) void kBigBounds(const Chunk* __restrict__ chunks,
                                                  const BigTask* __restrict__ tasks, TriSoA t,
                                                  const uint32_t* __restrict__ ind,
                                                  unsigned long long* __restrict__ bounds) {
    const Chunk ch = chunks[blockIdx.x];
    const uint32_t start = tasks[ch.task].start;
    unsigned long long v[6] = {kEmptyMin, kEmptyMin, kEmptyMin, kEmptyMax, kEmptyMax, kEmptyMax};
    for (uint32_t i = ch.begin + threadIdx.x; i < ch.end; i += 256u) {
        const uint32_t tri = ind[i], p = i - start;
        for (int k = 0; k < 3; ++k) {
            v[k] = umin64(v[k], packMin(t.mn[k][tri], p));
            v[3 + k] = umax64(v[3 + k], packMax(t.mx[k][tri], p));
        }
    }
    for (int j = 0; j < 3; ++j) v[j] = waveMin64(v[j]);
    for (int j = 3; j < 6; ++j) v[j] = waveMax64(v[j]);
    if ((threadIdx.x & 63u) == 0) {
        unsigned long long* b = bounds + 6 * size_t(ch.task);
        for (int j = 0; j < 3; ++j)
            if (v[j] != kEmptyMin) atomicMin(&b[j], v[j]);
        for (int j = 3; j < 6; ++j)
            if (v[j] != kEmptyMax) atomicMax(&b[j], v[j]);
    }
}

__global__ __launch_bounds__(64) void kBigDecode(const BigTask* __restrict__ tasks, uint32_t n,
                                                 const unsigned long long* __restrict__ bounds,
                                                 TriSoA t, const uint32_t* __restrict__ ind,
                                                 float* __restrict__ box) {
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    if (i >= 6 * n) return;
    const uint32_t task = i / 6, j = i % 6;
    box[i] = decodeKey(bounds[i], int(j), t, ind, tasks[task].start);
}

// Binning of one chunk (findBestSplit, Core/BVH.cu:214-256) into LDS, then
// the chunk's non-empty bins into the task's bins.  The LDS min / max is
// skipped when the bin already holds a key at least as good.
__global__ __launch_bounds__(256) void kBigBin(const Chunk* __restrict__ chunks,
                                               const BigTask* __restrict__ tasks,
                                               const float* __restrict__ box, TriSoA t,
                                               const uint32_t* __restrict__ ind, uint32_t nb,
                                               unsigned long long* __restrict__ bins,
                                               uint32_t* __restrict__ cnt) {
    __shared__ unsigned long long sBin[3 * kDeviceMaxBins * 6];
    __shared__ uint32_t sCnt[3 * kDeviceMaxBins];
    const Chunk ch = chunks[blockIdx.x];
    const uint32_t start = tasks[ch.task].start;
    for (uint32_t i = threadIdx.x; i < 3 * nb * 6; i += 256u)
        sBin[i] = (i % 6) < 3 ? kEmptyMin : kEmptyMax;
    for (uint32_t i = threadIdx.x; i < 3 * nb; i += 256u) sCnt[i] = 0;
    const float* bx = box + 6 * size_t(ch.task);
    float mn[3], scale[3];
    for (int k = 0; k < 3; ++k) {
        mn[k] = bx[k];
        scale[k] = float(nb) / (bx[3 + k] - bx[k]);
    }
    __syncthreads();
    const int NB = int(nb);
    for (uint32_t i = ch.begin + threadIdx.x; i < ch.end; i += 256u) {
        const uint32_t tri = ind[i], p = i - start;
        unsigned long long kv[6];
        for (int k = 0; k < 3; ++k) {
            kv[k] = packMin(t.mn[k][tri], p);
            kv[3 + k] = packMax(t.mx[k][tri], p);
        }
        for (int k = 0; k < 3; ++k) {
            int bi = min(NB - 1, truncX86((t.c[k][tri] - mn[k]) * scale[k]));
            bi = max(0, min(bi, NB - 1));
            unsigned long long* b = sBin + (size_t(k) * nb + bi) * 6;
            for (int j = 0; j < 3; ++j)
                if (kv[j] < b[j]) atomicMin(&b[j], kv[j]);
            for (int j = 3; j < 6; ++j)
                if (kv[j] > b[j]) atomicMax(&b[j], kv[j]);
            atomicAdd(&sCnt[size_t(k) * nb + bi], 1u);
        }
    }
    __syncthreads();
    unsigned long long* gb = bins + size_t(ch.task) * 3 * nb * 6;
    uint32_t* gc = cnt + size_t(ch.task) * 3 * nb;
    for (uint32_t i = threadIdx.x; i < 3 * nb; i += 256u) {
        if (sCnt[i] == 0) continue;
        atomicAdd(&gc[i], sCnt[i]);
        for (int j = 0; j < 3; ++j) atomicMin(&gb[6 * i + j], sBin[6 * i + j]);
        for (int j = 3; j < 6; ++j) atomicMax(&gb[6 * i + j], sBin[6 * i + j]);
    }
}

// The SAH scan of one task's bins (Core/BVH.cu:258-301): lane k scans axis k
// (the suffix boxes first, their areas and counts kept in LDS, then the
// prefix boxes with the cost of each split, first strict minimum in bin
// order); the caller takes the first strict minimum over the axes in order.
// sRsa / sRc: 3 x kDeviceMaxBins LDS scratch.
__device__ void sahScan(int lane, const unsigned long long* bins, const uint32_t* cnt, uint32_t nb,
                        const float* nbMn, const float* nbMx, const TriSoA& t,
                        const uint32_t* ind, uint32_t start, float* axisCost, float* axisSplit,
                        float* sRsa, uint32_t* sRc) {
    if (lane >= 3) return;
    const int k = lane, NB = int(nb);
    float* rsa = sRsa + k * kDeviceMaxBins;
    uint32_t* rcs = sRc + k * kDeviceMaxBins;
    float bmn[3] = {INFINITY, INFINITY, INFINITY}, bmx[3] = {-INFINITY, -INFINITY, -INFINITY};
    uint32_t cr = 0;
    for (int i = NB - 1; i > 0; --i) {  // rb[i - 1] = bins i .. NB - 1
        const unsigned long long* b = bins + (size_t(k) * nb + i) * 6;
        for (int j = 0; j < 3; ++j) {
            bmn[j] = fminS(bmn[j], decodeKey(b[j], j, t, ind, start));
            bmx[j] = fmaxS(bmx[j], decodeKey(b[3 + j], 3 + j, t, ind, start));
        }
        cr += cnt[size_t(k) * nb + i];
        rsa[i - 1] = boxArea(bmn, bmx);
        rcs[i - 1] = cr;
    }
    const float ext = nbMx[k] - nbMn[k];
    const float rootSA = boxArea(nbMn, nbMx);
    float best = INFINITY, split = 0.0f;
    for (int j = 0; j < 3; ++j) {
        bmn[j] = INFINITY;
        bmx[j] = -INFINITY;
    }
    int cc = 0;
    for (int i = 0; i < NB - 1; ++i) {  // lb[i] = bins 0 .. i
        const unsigned long long* b = bins + (size_t(k) * nb + i) * 6;
        for (int j = 0; j < 3; ++j) {
            bmn[j] = fminS(bmn[j], decodeKey(b[j], j, t, ind, start));
            bmx[j] = fmaxS(bmx[j], decodeKey(b[3 + j], 3 + j, t, ind, start));
        }
        cc += int(cnt[size_t(k) * nb + i]);
        const float splitPos = nbMn[k] + float(i + 1) * ext / float(NB);
        const float lsa = boxArea(bmn, bmx);
        const float cost = 1.0f + 1.0f * (float(cc) * lsa + float(int(rcs[i])) * rsa[i]) / rootSA;
        if (cost < best) {
            best = cost;
            split = splitPos;
        }
    }
    axisCost[k] = best;
    axisSplit[k] = split;
}

__global__ __launch_bounds__(64) void kBigSah(const uint32_t* __restrict__ active,
                                              const BigTask* __restrict__ tasks,
                                              const float* __restrict__ box,
                                              const unsigned long long* __restrict__ bins,
                                              const uint32_t* __restrict__ cnt, uint32_t nb,
                                              TriSoA t, const uint32_t* __restrict__ ind,
                                              uint32_t* __restrict__ axisOut,
                                              float* __restrict__ splitOut) {
    __shared__ float sCost[3], sSplit[3];
    __shared__ float sRsa[3 * kDeviceMaxBins];
    __shared__ uint32_t sRc[3 * kDeviceMaxBins];
    const uint32_t task = active[blockIdx.x];
    const float* bx = box + 6 * size_t(task);
    sahScan(int(threadIdx.x), bins + size_t(task) * 3 * nb * 6, cnt + size_t(task) * 3 * nb, nb,
            bx, bx + 3, t, ind, tasks[task].start, sCost, sSplit, sRsa, sRc);
    __syncthreads();
    if (threadIdx.x == 0) {
        float best = INFINITY, split = 0.0f;
        uint32_t axis = 0;
        for (uint32_t k = 0; k < 3; ++k)
            if (sCost[k] < best) {
                best = sCost[k];
                axis = k;
                split = sSplit[k];
            }
        axisOut[task] = axis;
        splitOut[task] = split;
    }
}

__global__ __launch_bounds__(256) void kBigPartCount(const Chunk* __restrict__ chunks,
                                                     const uint32_t* __restrict__ axis,
                                                     const float* __restrict__ split, TriSoA t,
                                                     const uint32_t* __restrict__ ind,
                                                     uint32_t* __restrict__ chunkTrue) {
    const Chunk ch = chunks[blockIdx.x];
    const float* c = t.c[axis[ch.task]];
    const float s = split[ch.task];
    uint32_t n = 0;
    for (uint32_t i = ch.begin + threadIdx.x; i < ch.end; i += 256u) n += c[ind[i]] < s ? 1u : 0u;
    n = waveSum(n);
    if ((threadIdx.x & 63u) == 0) atomicAdd(&chunkTrue[blockIdx.x], n);
}

// Per task (one thread): exclusive scan of its chunks' counts; the left count.
__global__ __launch_bounds__(64) void kBigPartScan(const uint32_t* __restrict__ firstChunk,
                                                   const uint32_t* __restrict__ nChunks,
                                                   uint32_t n, const uint32_t* __restrict__ chunkTrue,
                                                   uint32_t* __restrict__ chunkOff,
                                                   uint32_t* __restrict__ lcOut) {
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    if (i >= n) return;
    uint32_t run = 0;
    for (uint32_t c = firstChunk[i]; c < firstChunk[i] + nChunks[i]; ++c) {
        chunkOff[c] = run;
        run += chunkTrue[c];
    }
    lcOut[i] = run;
}

// Positions of the pairs the two-pointer partition swaps: a failing element
// left of lc with r failing elements before it goes to slot[start + r]; a
// passing element right of lc with r passing elements after it to
// slot[start + lc + r] (positions relative to the task).
__global__ __launch_bounds__(256) void kBigPartSlots(const Chunk* __restrict__ chunks,
                                                     const BigTask* __restrict__ tasks,
                                                     const uint32_t* __restrict__ axis,
                                                     const float* __restrict__ split, TriSoA t,
                                                     const uint32_t* __restrict__ ind,
                                                     const uint32_t* __restrict__ chunkOff,
                                                     const uint32_t* __restrict__ taskLc,
                                                     uint32_t* __restrict__ slot,
                                                     uint32_t* __restrict__ taskM) {
    __shared__ uint32_t sWave[4];
    const Chunk ch = chunks[blockIdx.x];
    const uint32_t start = tasks[ch.task].start, lc = taskLc[ch.task];
    const float* c = t.c[axis[ch.task]];
    const float s = split[ch.task];
    uint32_t run = chunkOff[blockIdx.x], m = 0;
    for (uint32_t base = ch.begin; base < ch.end; base += 256u) {
        const uint32_t i = base + threadIdx.x;
        const bool in = i < ch.end;
        const bool pred = in && c[ind[i]] < s;
        uint32_t tot;
        const uint32_t before = run + blockPrefix(pred, sWave, &tot);
        if (in) {
            const uint32_t p = i - start;
            if (p < lc && !pred) {
                slot[start + (p - before)] = p;
                ++m;
            } else if (p >= lc && pred) {
                slot[start + lc + (lc - 1u - before)] = p;
            }
        }
        run += tot;
    }
    m = waveSum(m);
    if ((threadIdx.x & 63u) == 0 && m) atomicAdd(&taskM[ch.task], m);
}

__global__ __launch_bounds__(256) void kBigSwap(const uint32_t* __restrict__ active,
                                                const BigTask* __restrict__ tasks,
                                                const uint32_t* __restrict__ taskLc,
                                                const uint32_t* __restrict__ taskM,
                                                const uint32_t* __restrict__ slot,
                                                uint32_t* __restrict__ ind) {
    const uint32_t task = active[blockIdx.y];
    const uint32_t start = tasks[task].start, lc = taskLc[task], m = taskM[task];
    for (uint32_t r = blockIdx.x * 256u + threadIdx.x; r < m; r += gridDim.x * 256u) {
        const uint32_t a = start + slot[start + r], b = start + slot[start + lc + r];
        const uint32_t x = ind[a];
        ind[a] = ind[b];
        ind[b] = x;
    }
}

// ---- small tasks: one workgroup per subtree -------------------------------------
struct StackEntry {
    uint32_t s, n, depth, parent;  // parent: (device node << 1) | side, or ~0u for the task root
};

__global__ __launch_bounds__(256) void kSmallBuild(const SmallTask* __restrict__ tasks, TriSoA t,
                                                   uint32_t* __restrict__ ind,
                                                   DNode* __restrict__ nodes,
                                                   uint32_t* __restrict__ counter,
                                                   uint32_t* __restrict__ rootOut, uint32_t nb,
                                                   uint32_t maxLeaf, uint32_t maxDepth) {
    __shared__ uint32_t sInd[kSmallMax];
    __shared__ uint32_t sSlot[kSmallMax];
    __shared__ unsigned long long sBin[3 * kDeviceMaxBins * 6];
    __shared__ uint32_t sCnt[3 * kDeviceMaxBins];
    __shared__ unsigned long long sNb[6];
    __shared__ float sBox[6], sCost[3], sSplitK[3];
    __shared__ float sRsa[3 * kDeviceMaxBins];
    __shared__ uint32_t sRc[3 * kDeviceMaxBins];
    __shared__ StackEntry sStack[kStackMax];
    __shared__ uint32_t sTop, sNode, sLc, sM, sAxis, sWave[4];
    __shared__ float sSplit;
    __shared__ StackEntry sCur;
    const SmallTask task = tasks[blockIdx.x];
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < task.count; i += 256u) sInd[i] = ind[task.start + i];
    if (tid == 0) {
        sStack[0] = StackEntry{0u, task.count, task.depth, ~0u};
        sTop = 1;
    }
    __syncthreads();
    const int NB = int(nb);
    while (true) {
        if (tid == 0) sCur = sStack[--sTop];
        if (tid < 6) sNb[tid] = tid < 3 ? kEmptyMin : kEmptyMax;
        __syncthreads();
        const StackEntry e = sCur;
        // node bounds
        unsigned long long v[6] = {kEmptyMin, kEmptyMin, kEmptyMin, kEmptyMax, kEmptyMax, kEmptyMax};
        for (uint32_t i = tid; i < e.n; i += 256u) {
            const uint32_t tri = sInd[e.s + i];
            for (int k = 0; k < 3; ++k) {
                v[k] = umin64(v[k], packMin(t.mn[k][tri], i));
                v[3 + k] = umax64(v[3 + k], packMax(t.mx[k][tri], i));
            }
        }
        for (int j = 0; j < 3; ++j) v[j] = waveMin64(v[j]);
        for (int j = 3; j < 6; ++j) v[j] = waveMax64(v[j]);
        if ((tid & 63u) == 0) {
            for (int j = 0; j < 3; ++j) atomicMin(&sNb[j], v[j]);
            for (int j = 3; j < 6; ++j) atomicMax(&sNb[j], v[j]);
        }
        __syncthreads();
        if (tid < 6) {
            const unsigned long long key = sNb[tid];
            const bool isMax = tid >= 3;
            float val;
            if (key == (isMax ? kEmptyMax : kEmptyMin)) {
                val = isMax ? -INFINITY : INFINITY;
            } else {
                const uint32_t low = uint32_t(key), pos = isMax ? 0xFFFFFFFFu - low : low;
                const uint32_t tri = sInd[e.s + pos];
                val = isMax ? t.mx[tid - 3][tri] : t.mn[tid][tri];
            }
            sBox[tid] = val;
        }
        if (tid == 0) {
            const uint32_t id = atomicAdd(counter, 1u);
            sNode = id;
            if (e.parent == ~0u) {
                rootOut[blockIdx.x] = id;
            } else if (e.parent & 1u) {
                nodes[e.parent >> 1].b = id;
            } else {
                nodes[e.parent >> 1].a = id;
            }
        }
        __syncthreads();
        const uint32_t id = sNode;
        bool leaf = e.n <= maxLeaf || e.depth >= maxDepth;
        if (!leaf) {
            // binning
            for (uint32_t i = tid; i < 3 * nb * 6; i += 256u)
                sBin[i] = (i % 6) < 3 ? kEmptyMin : kEmptyMax;
            for (uint32_t i = tid; i < 3 * nb; i += 256u) sCnt[i] = 0;
            __syncthreads();
            float mn[3], scale[3];
            for (int k = 0; k < 3; ++k) {
                mn[k] = sBox[k];
                scale[k] = float(nb) / (sBox[3 + k] - sBox[k]);
            }
            for (uint32_t i = tid; i < e.n; i += 256u) {
                const uint32_t tri = sInd[e.s + i];
                unsigned long long kv[6];
                for (int k = 0; k < 3; ++k) {
                    kv[k] = packMin(t.mn[k][tri], i);
                    kv[3 + k] = packMax(t.mx[k][tri], i);
                }
                for (int k = 0; k < 3; ++k) {
                    int bi = min(NB - 1, truncX86((t.c[k][tri] - mn[k]) * scale[k]));
                    bi = max(0, min(bi, NB - 1));
                    unsigned long long* b = sBin + (size_t(k) * nb + bi) * 6;
                    for (int j = 0; j < 3; ++j)
                        if (kv[j] < b[j]) atomicMin(&b[j], kv[j]);
                    for (int j = 3; j < 6; ++j)
                        if (kv[j] > b[j]) atomicMax(&b[j], kv[j]);
                    atomicAdd(&sCnt[size_t(k) * nb + bi], 1u);
                }
            }
            __syncthreads();
            // bins decode through sInd: positions are relative to e.s
            if (tid < 64)
                sahScan(int(tid), sBin, sCnt, nb, sBox, sBox + 3, t, sInd, e.s, sCost, sSplitK,
                        sRsa, sRc);
            __syncthreads();
            if (tid == 0) {
                float best = INFINITY, split = 0.0f;
                uint32_t axis = 0;
                for (uint32_t k = 0; k < 3; ++k)
                    if (sCost[k] < best) {
                        best = sCost[k];
                        axis = k;
                        split = sSplitK[k];
                    }
                sAxis = axis;
                sSplit = split;
                sLc = 0;
                sM = 0;
            }
            __syncthreads();
            const float* c = t.c[sAxis];
            const float s = sSplit;
            uint32_t cntTrue = 0;
            for (uint32_t i = tid; i < e.n; i += 256u) cntTrue += c[sInd[e.s + i]] < s ? 1u : 0u;
            cntTrue = waveSum(cntTrue);
            if ((tid & 63u) == 0 && cntTrue) atomicAdd(&sLc, cntTrue);
            __syncthreads();
            const uint32_t lc = sLc;
            if (lc == 0 || lc == e.n) {
                leaf = true;
            } else {
                uint32_t run = 0, m = 0;
                for (uint32_t base = 0; base < e.n; base += 256u) {
                    const uint32_t p = base + tid;
                    const bool in = p < e.n;
                    const bool pred = in && c[sInd[e.s + p]] < s;
                    uint32_t tot;
                    const uint32_t before = run + blockPrefix(pred, sWave, &tot);
                    if (in) {
                        if (p < lc && !pred) {
                            sSlot[e.s + (p - before)] = p;
                            ++m;
                        } else if (p >= lc && pred) {
                            sSlot[e.s + lc + (lc - 1u - before)] = p;
                        }
                    }
                    run += tot;
                }
                m = waveSum(m);
                if ((tid & 63u) == 0 && m) atomicAdd(&sM, m);
                __syncthreads();
                for (uint32_t r = tid; r < sM; r += 256u) {
                    const uint32_t a = e.s + sSlot[e.s + r], b = e.s + sSlot[e.s + lc + r];
                    const uint32_t x = sInd[a];
                    sInd[a] = sInd[b];
                    sInd[b] = x;
                }
                if (tid == 0) {
                    DNode& o = nodes[id];
                    for (int k = 0; k < 3; ++k) {
                        o.mn[k] = sBox[k];
                        o.mx[k] = sBox[3 + k];
                    }
                    o.leaf = 0;
                    o.a = o.b = 0;
                    // right pushed first: the left subtree is built first
                    sStack[sTop++] = StackEntry{e.s + lc, e.n - lc, e.depth + 1, (id << 1) | 1u};
                    sStack[sTop++] = StackEntry{e.s, lc, e.depth + 1, id << 1};
                }
            }
        }
        if (leaf && tid == 0) {
            DNode& o = nodes[id];
            for (int k = 0; k < 3; ++k) {
                o.mn[k] = sBox[k];
                o.mx[k] = sBox[3 + k];
            }
            o.leaf = 1;
            o.a = task.start + e.s;
            o.b = e.n;
        }
        __syncthreads();
        const uint32_t top = sTop;  // read by every thread before thread 0 pops again
        __syncthreads();
        if (top == 0) break;
    }
    for (uint32_t i = tid; i < task.count; i += 256u) ind[task.start + i] = sInd[i];
}

// ---- host side -------------------------------------------------------------------
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t n) {
        if (p) (void)hipFree(p);
        p = nullptr;
        return n ? hipMalloc(&p, n) : hipSuccess;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

#define BVHD_TRY(x)                       \
    do {                                  \
        hipError_t e_ = (x);              \
        if (e_ != hipSuccess) return e_;  \
    } while (0)

struct HNode {
    float mn[3], mx[3];
    uint32_t leaf, a, b;  // internal: child refs (kDevBit | device node, or host node)
};

struct Task {
    uint32_t start, count, depth;
    uint32_t parent, side;  // parent host node (~0u: root)
};

}  // namespace

hipError_t buildBvhDevice(const float* dPos, const uint32_t* dIdx, uint32_t ntri,
                          const BuildParams& p, hipStream_t st, BvhResult& out) {
    if (p.bins < 2 || p.bins > kDeviceMaxBins || p.maxDepth > kDeviceMaxDepth)
        return hipErrorInvalidValue;
    const auto t0 = std::chrono::steady_clock::now();
    out = BvhResult{};
    const uint32_t nb = p.bins;
    const size_t N = ntri;
    DevBuf soa, indB, slotB, nodeB, cntB, rootB;
    TriSoA t;
    BVHD_TRY(soa.alloc(9 * std::max<size_t>(N, 1) * sizeof(float)));
    for (int k = 0; k < 3; ++k) {
        t.c[k] = soa.as<float>() + size_t(k) * N;
        t.mn[k] = soa.as<float>() + size_t(3 + k) * N;
        t.mx[k] = soa.as<float>() + size_t(6 + k) * N;
    }
    BVHD_TRY(indB.alloc(std::max<size_t>(N, 1) * 4));
    BVHD_TRY(slotB.alloc(std::max<size_t>(N, 1) * 4));
    BVHD_TRY(nodeB.alloc((2 * N + 1) * sizeof(DNode)));
    BVHD_TRY(cntB.alloc(64));
    uint32_t* ind = indB.as<uint32_t>();
    if (N) {
        hipLaunchKernelGGL(kPrep, dim3(uint32_t(std::min<size_t>((N + 255) / 256, 65536))),
                           dim3(256), 0, st, dPos, dIdx, ntri, t, ind);
        BVHD_TRY(hipGetLastError());
    }
    BVHD_TRY(hipMemsetAsync(cntB.p, 0, 64, st));

    std::vector<HNode> hn;
    std::vector<Task> level{Task{0u, ntri, 0u, ~0u, 0u}}, next;
    std::vector<Task> small;
    auto link = [&](const Task& tk, uint32_t ref) {
        if (tk.parent == ~0u) return;
        (tk.side ? hn[tk.parent].b : hn[tk.parent].a) = ref;
    };
    uint32_t rootRef = ~0u;
    DevBuf tasksB, chunksB, boundsB, boxB, binsB, bcntB, actB, axisB, splitB, ctB, coB, lcB, mB,
        fcB, ncB;
    while (!level.empty()) {
        std::vector<Task> big;
        for (const Task& tk : level)
            (tk.count > kSmallMax ? big : small).push_back(tk);
        level.clear();
        if (big.empty()) break;
        const uint32_t nt = uint32_t(big.size());
        std::vector<BigTask> bt(nt);
        std::vector<Chunk> chunks;
        for (uint32_t i = 0; i < nt; ++i) {
            bt[i] = BigTask{big[i].start, big[i].count, big[i].depth, 0u};
            for (uint32_t b = big[i].start; b < big[i].start + big[i].count; b += kChunk)
                chunks.push_back(Chunk{i, b, std::min(b + kChunk, big[i].start + big[i].count), 0u});
        }
        BVHD_TRY(tasksB.alloc(nt * sizeof(BigTask)));
        BVHD_TRY(chunksB.alloc(chunks.size() * sizeof(Chunk)));
        BVHD_TRY(boundsB.alloc(size_t(nt) * 6 * 8));
        BVHD_TRY(boxB.alloc(size_t(nt) * 6 * 4));
        BVHD_TRY(hipMemcpyAsync(tasksB.p, bt.data(), nt * sizeof(BigTask), hipMemcpyHostToDevice, st));
        BVHD_TRY(hipMemcpyAsync(chunksB.p, chunks.data(), chunks.size() * sizeof(Chunk),
                                hipMemcpyHostToDevice, st));
        {
            std::vector<unsigned long long> init(size_t(nt) * 6);
            for (size_t i = 0; i < init.size(); ++i) init[i] = (i % 6) < 3 ? kEmptyMin : kEmptyMax;
            BVHD_TRY(hipMemcpyAsync(boundsB.p, init.data(), init.size() * 8,
                                    hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(kBigBounds, dim3(uint32_t(chunks.size())), dim3(256), 0, st,
                               chunksB.as<Chunk>(), tasksB.as<BigTask>(), t, ind,
                               boundsB.as<unsigned long long>());
            hipLaunchKernelGGL(kBigDecode, dim3((6 * nt + 63) / 64), dim3(64), 0, st,
                               tasksB.as<BigTask>(), nt, boundsB.as<unsigned long long>(), t, ind,
                               boxB.as<float>());
            BVHD_TRY(hipGetLastError());
        }
        std::vector<float> box(size_t(nt) * 6);
        BVHD_TRY(hipMemcpyAsync(box.data(), boxB.p, box.size() * 4, hipMemcpyDeviceToHost, st));
        BVHD_TRY(hipStreamSynchronize(st));
        std::vector<uint32_t> nodeOf(nt), active;
        for (uint32_t i = 0; i < nt; ++i) {
            HNode h;
            std::memcpy(h.mn, &box[6 * i], 12);
            std::memcpy(h.mx, &box[6 * i + 3], 12);
            const bool leaf = big[i].count <= p.maxLeaf || big[i].depth >= p.maxDepth;
            h.leaf = leaf ? 1u : 0u;
            h.a = big[i].start;
            h.b = big[i].count;
            nodeOf[i] = uint32_t(hn.size());
            hn.push_back(h);
            link(big[i], nodeOf[i]);
            if (big[i].parent == ~0u) rootRef = nodeOf[i];
            if (!leaf) active.push_back(i);
        }
        if (active.empty()) continue;
        const uint32_t na = uint32_t(active.size());
        std::vector<Chunk> ach;
        std::vector<uint32_t> firstChunk(na), nChunks(na);
        for (uint32_t j = 0; j < na; ++j) {
            const Task& tk = big[active[j]];
            firstChunk[j] = uint32_t(ach.size());
            for (uint32_t b = tk.start; b < tk.start + tk.count; b += kChunk)
                ach.push_back(Chunk{active[j], b, std::min(b + kChunk, tk.start + tk.count), 0u});
            nChunks[j] = uint32_t(ach.size()) - firstChunk[j];
        }
        const size_t nbins = size_t(nt) * 3 * nb;
        BVHD_TRY(binsB.alloc(nbins * 6 * 8));
        BVHD_TRY(bcntB.alloc(nbins * 4));
        BVHD_TRY(actB.alloc(na * 4));
        BVHD_TRY(axisB.alloc(nt * 4));
        BVHD_TRY(splitB.alloc(nt * 4));
        BVHD_TRY(ctB.alloc(ach.size() * 4));
        BVHD_TRY(coB.alloc(ach.size() * 4));
        BVHD_TRY(lcB.alloc(na * 4));
        BVHD_TRY(mB.alloc(nt * 4));
        BVHD_TRY(fcB.alloc(na * 4));
        BVHD_TRY(ncB.alloc(na * 4));
        {
            std::vector<unsigned long long> init(nbins * 6);
            for (size_t i = 0; i < init.size(); ++i) init[i] = (i % 6) < 3 ? kEmptyMin : kEmptyMax;
            BVHD_TRY(hipMemcpyAsync(binsB.p, init.data(), init.size() * 8, hipMemcpyHostToDevice,
                                    st));
        }
        BVHD_TRY(hipMemsetAsync(bcntB.p, 0, nbins * 4, st));
        BVHD_TRY(hipMemsetAsync(ctB.p, 0, ach.size() * 4, st));
        BVHD_TRY(hipMemsetAsync(mB.p, 0, nt * 4, st));
        BVHD_TRY(hipMemcpyAsync(actB.p, active.data(), na * 4, hipMemcpyHostToDevice, st));
        BVHD_TRY(hipMemcpyAsync(fcB.p, firstChunk.data(), na * 4, hipMemcpyHostToDevice, st));
        BVHD_TRY(hipMemcpyAsync(ncB.p, nChunks.data(), na * 4, hipMemcpyHostToDevice, st));
        // ach is the active tasks' chunks; reuse chunksB for it
        BVHD_TRY(chunksB.alloc(ach.size() * sizeof(Chunk)));
        BVHD_TRY(hipMemcpyAsync(chunksB.p, ach.data(), ach.size() * sizeof(Chunk),
                                hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(kBigBin, dim3(uint32_t(ach.size())), dim3(256), 0, st,
                           chunksB.as<Chunk>(), tasksB.as<BigTask>(), boxB.as<float>(), t, ind, nb,
                           binsB.as<unsigned long long>(), bcntB.as<uint32_t>());
        hipLaunchKernelGGL(kBigSah, dim3(na), dim3(64), 0, st, actB.as<uint32_t>(),
                           tasksB.as<BigTask>(), boxB.as<float>(), binsB.as<unsigned long long>(),
                           bcntB.as<uint32_t>(), nb, t, ind, axisB.as<uint32_t>(),
                           splitB.as<float>());
        hipLaunchKernelGGL(kBigPartCount, dim3(uint32_t(ach.size())), dim3(256), 0, st,
                           chunksB.as<Chunk>(), axisB.as<uint32_t>(), splitB.as<float>(), t, ind,
                           ctB.as<uint32_t>());
        hipLaunchKernelGGL(kBigPartScan, dim3((na + 63) / 64), dim3(64), 0, st,
                           fcB.as<uint32_t>(), ncB.as<uint32_t>(), na, ctB.as<uint32_t>(),
                           coB.as<uint32_t>(), lcB.as<uint32_t>());
        BVHD_TRY(hipGetLastError());
        std::vector<uint32_t> lcA(na);
        BVHD_TRY(hipMemcpyAsync(lcA.data(), lcB.p, na * 4, hipMemcpyDeviceToHost, st));
        BVHD_TRY(hipStreamSynchronize(st));
        // per-task left counts indexed by task (the slot / swap kernels read lc by task)
        std::vector<uint32_t> lcT(nt, 0);
        std::vector<uint32_t> split;  // active tasks that really split
        for (uint32_t j = 0; j < na; ++j) {
            const uint32_t i = active[j];
            lcT[i] = lcA[j];
            HNode& h = hn[nodeOf[i]];
            if (lcA[j] == 0 || lcA[j] == big[i].count) {
                h.leaf = 1;  // degenerate split (Core/BVH.cu:186-195)
                continue;
            }
            split.push_back(j);
        }
        if (!split.empty()) {
            std::vector<Chunk> sch;
            std::vector<uint32_t> sOff, sTask;
            for (uint32_t j : split) {
                for (uint32_t c = firstChunk[j]; c < firstChunk[j] + nChunks[j]; ++c) {
                    sch.push_back(ach[c]);
                    sOff.push_back(c);
                }
                sTask.push_back(active[j]);
            }
            DevBuf lcTB, schB, soB, stB, coS;
            BVHD_TRY(lcTB.alloc(nt * 4));
            BVHD_TRY(hipMemcpyAsync(lcTB.p, lcT.data(), nt * 4, hipMemcpyHostToDevice, st));
            // chunk offsets of the splitting tasks' chunks, in sch order
            std::vector<uint32_t> co(ach.size());
            BVHD_TRY(hipMemcpyAsync(co.data(), coB.p, ach.size() * 4, hipMemcpyDeviceToHost, st));
            BVHD_TRY(hipStreamSynchronize(st));
            std::vector<uint32_t> coSel(sch.size());
            for (size_t q = 0; q < sch.size(); ++q) coSel[q] = co[sOff[q]];
            BVHD_TRY(schB.alloc(sch.size() * sizeof(Chunk)));
            BVHD_TRY(coS.alloc(sch.size() * 4));
            BVHD_TRY(stB.alloc(sTask.size() * 4));
            BVHD_TRY(hipMemcpyAsync(schB.p, sch.data(), sch.size() * sizeof(Chunk),
                                    hipMemcpyHostToDevice, st));
            BVHD_TRY(hipMemcpyAsync(coS.p, coSel.data(), sch.size() * 4, hipMemcpyHostToDevice,
                                    st));
            BVHD_TRY(hipMemcpyAsync(stB.p, sTask.data(), sTask.size() * 4, hipMemcpyHostToDevice,
                                    st));
            hipLaunchKernelGGL(kBigPartSlots, dim3(uint32_t(sch.size())), dim3(256), 0, st,
                               schB.as<Chunk>(), tasksB.as<BigTask>(), axisB.as<uint32_t>(),
                               splitB.as<float>(), t, ind, coS.as<uint32_t>(),
                               lcTB.as<uint32_t>(), slotB.as<uint32_t>(), mB.as<uint32_t>());
            uint32_t maxCount = 0;
            for (uint32_t i : sTask) maxCount = std::max(maxCount, big[i].count);
            const uint32_t gx = std::min<uint32_t>((maxCount / 2 + 255) / 256, 256);
            hipLaunchKernelGGL(kBigSwap, dim3(std::max(gx, 1u), uint32_t(sTask.size())), dim3(256),
                               0, st, stB.as<uint32_t>(), tasksB.as<BigTask>(),
                               lcTB.as<uint32_t>(), mB.as<uint32_t>(), slotB.as<uint32_t>(), ind);
            BVHD_TRY(hipGetLastError());
            BVHD_TRY(hipStreamSynchronize(st));  // the temporaries above go out of scope
        }
        for (uint32_t j : split) {
            const uint32_t i = active[j];
            const Task& tk = big[i];
            const uint32_t lc = lcT[i];
            level.push_back(Task{tk.start, lc, tk.depth + 1, nodeOf[i], 0u});
            level.push_back(Task{tk.start + lc, tk.count - lc, tk.depth + 1, nodeOf[i], 1u});
        }
    }
    // every task of at most kSmallMax triangles, one workgroup each
    uint32_t devCount = 0;
    std::vector<uint32_t> roots(small.size());
    if (!small.empty()) {
        std::vector<SmallTask> stk(small.size());
        for (size_t i = 0; i < small.size(); ++i)
            stk[i] = SmallTask{small[i].start, small[i].count, small[i].depth, 0u};
        DevBuf sB;
        BVHD_TRY(sB.alloc(stk.size() * sizeof(SmallTask)));
        BVHD_TRY(rootB.alloc(stk.size() * 4));
        BVHD_TRY(hipMemcpyAsync(sB.p, stk.data(), stk.size() * sizeof(SmallTask),
                                hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(kSmallBuild, dim3(uint32_t(stk.size())), dim3(256), 0, st,
                           sB.as<SmallTask>(), t, ind, nodeB.as<DNode>(), cntB.as<uint32_t>(),
                           rootB.as<uint32_t>(), nb, p.maxLeaf, p.maxDepth);
        BVHD_TRY(hipGetLastError());
        BVHD_TRY(hipMemcpyAsync(roots.data(), rootB.p, roots.size() * 4, hipMemcpyDeviceToHost,
                                st));
        BVHD_TRY(hipMemcpyAsync(&devCount, cntB.p, 4, hipMemcpyDeviceToHost, st));
        BVHD_TRY(hipStreamSynchronize(st));
    }
    std::vector<DNode> dn(devCount);
    if (devCount)
        BVHD_TRY(hipMemcpyAsync(dn.data(), nodeB.p, size_t(devCount) * sizeof(DNode),
                                hipMemcpyDeviceToHost, st));
    out.primOrder.resize(N);
    if (N) BVHD_TRY(hipMemcpyAsync(out.primOrder.data(), ind, N * 4, hipMemcpyDeviceToHost, st));
    BVHD_TRY(hipStreamSynchronize(st));
    for (size_t i = 0; i < small.size(); ++i) {
        link(small[i], kDevBit | roots[i]);
        if (small[i].parent == ~0u) rootRef = kDevBit | roots[i];
    }
    // Post-order emission over both node sets (left, right, node; root last),
    // iterative: (ref, level, children done).
    out.nodes.reserve(hn.size() + devCount);
    struct Frame {
        uint32_t ref, level, state, left;
    };
    std::vector<Frame> fs{Frame{rootRef, 0u, 0u, 0u}};
    uint32_t lastEmitted = 0;
    while (!fs.empty()) {
        Frame& f = fs.back();
        const bool dev = (f.ref & kDevBit) != 0;
        const uint32_t id = f.ref & ~kDevBit;
        const float* mn = dev ? dn[id].mn : hn[id].mn;
        const float* mx = dev ? dn[id].mx : hn[id].mx;
        const uint32_t leaf = dev ? dn[id].leaf : hn[id].leaf;
        const uint32_t a = dev ? dn[id].a : hn[id].a, b = dev ? dn[id].b : hn[id].b;
        out.levels = std::max(out.levels, f.level);
        crh_bvh_node_ref r;
        if (leaf) {
            r.tag = 0;
            r.a = a;
            r.b = b;
            out.maxLeaf = std::max(out.maxLeaf, b);
        } else if (f.state == 0) {
            f.state = 1;
            // device children refer to device nodes
            const uint32_t cref = dev ? (kDevBit | a) : a;
            fs.push_back(Frame{cref, f.level + 1, 0u, 0u});
            continue;
        } else if (f.state == 1) {
            f.state = 2;
            f.left = lastEmitted;
            const uint32_t cref = dev ? (kDevBit | b) : b;
            fs.push_back(Frame{cref, f.level + 1, 0u, 0u});
            continue;
        } else {
            r.tag = 1;
            r.a = f.left;
            r.b = lastEmitted;
            out.internalCount++;
            out.maxDepth = std::max(out.maxDepth, f.level);
        }
        for (int k = 0; k < 3; ++k) {
            r.pmin[k] = mn[k];
            r.pmax[k] = mx[k];
        }
        out.nodes.push_back(r);
        lastEmitted = uint32_t(out.nodes.size() - 1);
        fs.pop_back();
    }
    out.root = lastEmitted;
    const crh_bvh_node_ref& rn = out.nodes[out.root];
    for (int k = 0; k < 3; ++k) {
        out.rootMin[k] = rn.pmin[k];
        out.rootMax[k] = rn.pmax[k];
    }
    out.buildMs =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return hipSuccess;
}

}  // namespace crh
