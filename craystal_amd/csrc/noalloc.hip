// Allocation-free closest-hit query: the variant of crh_trace_closest_device
// whose scratch (the persistent kernel's work dispenser counter, the per-ray
// node / triangle counts of the counting kernel and their totals) lives in
// one file-scope device array instead of memory the library allocates.
//
// The array `scratch` holds kNoAllocHeadBytes of header (dispenser counter,
// then the two work totals) followed by the per-ray counts of one chunk of
// at most kNoAllocChunkRays rays; longer batches run chunk by chunk.  Each
// GPU holds its own copy of the array, at its own address, found once per
// device by hipGetSymbolAddress (noallocScratchInit, called from
// crh_context_create with the device current); host code passes that
// address to the kernels, never the array's name.  Every call on a GPU
// shares that GPU's array and overwrites it: calls that overlap on one GPU
// (two host threads, or two contexts on one device) race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "kernels.h"
#include "noalloc.h"

constexpr size_t kScratchBytes =
    crh::kNoAllocHeadBytes + size_t(crh::kNoAllocChunkRays) * 2 * sizeof(uint32_t);

// 8 MiB + 256 B per GPU, held from module load until the process ends.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
void* gScratch[kMaxDevices] = {};

// Header layout (bytes): 0 dispenser counter (u32), 64 node total (u64),
// 72 triangle total (u64).
constexpr size_t kOffCounter = 0, kOffTotals = 64;

__global__ __launch_bounds__(64) void kZeroTotals(unsigned long long* t) {
    if (threadIdx.x < 2) t[threadIdx.x] = 0ull;
}

// Adds the per-ray (nodes, triangles) pairs of work[0, n) into t[0], t[1]:
// a block sums its share, then one atomic per block and total.
__global__ __launch_bounds__(256) void kSumWork(const uint32_t* __restrict__ work, uint32_t n,
                                                unsigned long long* __restrict__ t) {
    __shared__ unsigned long long part[2][4];
    unsigned long long a = 0, b = 0;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        a += work[2 * size_t(i)];
        b += work[2 * size_t(i) + 1];
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
    }
    const uint32_t wave = threadIdx.x / 64u, lane = threadIdx.x % 64u;
    if (lane == 0) {
        part[0][wave] = a;
        part[1][wave] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&t[0], part[0][0] + part[0][1] + part[0][2] + part[0][3]);
        atomicAdd(&t[1], part[1][0] + part[1][1] + part[1][2] + part[1][3]);
    }
}

}  // namespace

namespace crh {

hipError_t noallocScratchInit(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(gScratchMutex);
    if (gScratch[device]) return hipSuccess;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    gScratch[device] = p;
    return hipSuccess;
}

size_t noallocScratchBytes() { return kScratchBytes; }

hipError_t traceClosestNoAlloc(int device, const SceneView& s, const crh_ray* rays, uint64_t n,
                               crh_hit* hits, bool countWork, const Tuning& t, hipStream_t st,
                               unsigned long long totals[2]) {
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(gScratchMutex);
        if (device >= 0 && device < kMaxDevices) base = gScratch[device];
    }
    if (!base) return hipErrorNotInitialized;
    unsigned char* sb = static_cast<unsigned char*>(base);
    auto* counter = reinterpret_cast<uint32_t*>(sb + kOffCounter);
    auto* tot = reinterpret_cast<unsigned long long*>(sb + kOffTotals);
    auto* work = reinterpret_cast<uint32_t*>(sb + kNoAllocHeadBytes);
    if (!countWork) {
        // The persistent kernel resets the counter itself before it starts.
        return launchTraceClosest(s, rays, n, hits, nullptr, t, counter, st);
    }
    hipLaunchKernelGGL(kZeroTotals, dim3(1), dim3(64), 0, st, tot);
    for (uint64_t off = 0; off < n; off += kNoAllocChunkRays) {
        const uint32_t m = uint32_t(std::min<uint64_t>(kNoAllocChunkRays, n - off));
        hipError_t e = launchTraceClosest(s, rays + off, m, hits + off, work, t, nullptr, st);
        if (e != hipSuccess) return e;
        const uint32_t blocks = std::min<uint32_t>(1024u, (m + 255u) / 256u);
        hipLaunchKernelGGL(kSumWork, dim3(blocks), dim3(256), 0, st, work, m, tot);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(totals, tot, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(st);
}

}  // namespace crh
