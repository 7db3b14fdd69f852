// Learned path guiding: a small MLP maps (position, normal, outgoing
// direction) to an 8-lobe von Mises-Fisher mixture over incident directions.
// Layer shapes, the feature encoding and the head are fixed here and shared
// by every kernel (guiding.hip) and the host code.
#pragma once
#include <cstdint>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define CRH_GHD constexpr __host__ __device__ __forceinline__
#else
#define CRH_GHD constexpr inline
#endif

namespace crh {

constexpr int kGIn = 64;      // padded input features
constexpr int kGWidth = 64;   // hidden width
constexpr int kGOut = 48;     // output rows (40 used: 8 lobes x 5)
constexpr int kGLobes = 8;
constexpr int kGLayers = 5;   // 4 hidden layers + output layer
constexpr int kGFreq = 5;     // frequency levels per axis
constexpr float kGKappaMin = 0.01f;

CRH_GHD int gLayerK(int l) { return l < kGLayers ? kGIn : 0; }
CRH_GHD int gLayerN(int l) { return l < kGLayers - 1 ? kGWidth : kGOut; }
// Flat fp32 parameter vector: for each layer W (K x N, row-major) then b (N).
CRH_GHD int gWOff(int l) {
    int off = 0;
    for (int i = 0; i < l; ++i) off += gLayerK(i) * gLayerN(i) + gLayerN(i);
    return off;
}
CRH_GHD int gBOff(int l) { return gWOff(l) + gLayerK(l) * gLayerN(l); }
constexpr int kGParams = 4 * (64 * 64 + 64) + (64 * 48 + 48);  // 19,760

// bf16 copies of the weights in MFMA B-operand order (v_mfma_f32_16x16x32_bf16:
// lane l supplies B[8*(l/16)+i][l%16], i < 8, of each 32 x 16 block).
// W_l (64 x N_l) and, for the backward pass, W_l^T zero-padded to 64 x 64.
CRH_GHD int gQOff(int l) { return l * 64 * 64; }  // elements, both copies
constexpr int kGQElems = kGLayers * 64 * 64;
// Element (k, n) of a K x N matrix (N multiple of 16) in operand order.
CRH_GHD int gSwz(int k, int n, int N) {
    int s = k >> 5, g = (k & 31) >> 3, i = k & 7, t = n >> 4;
    int lane = g * 16 + (n & 15);
    return ((s * (N >> 4) + t) * 64 + lane) * 8 + i;
}

// One training sample: shading point, normal, outgoing direction and the
// sampled incident direction; wi.w = target / sampling pdf.
struct alignas(16) GuideSample {
    float pos[4];
    float nrm[4];
    float wo[4];
    float wi[4];
};

}  // namespace crh
