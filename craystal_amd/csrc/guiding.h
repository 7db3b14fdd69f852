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
constexpr int kGWidth = 64;   // hidden width of the default network
constexpr int kGOut = 48;     // output rows (40 used: 8 lobes x 5)
constexpr int kGLobes = 8;
constexpr int kGLayers = 5;   // 4 hidden layers + output layer
constexpr int kGFreq = 5;     // frequency levels per axis
constexpr float kGKappaMin = 0.01f;

// Element (k, n) of a K x N matrix (N a multiple of 16) in MFMA B-operand
// order (v_mfma_f32_16x16x32_bf16: lane l supplies B[8*(l/16)+i][l%16],
// i < 8, of each 32 x 16 block).
CRH_GHD int gSwz(int k, int n, int N) {
    int s = k >> 5, g = (k & 31) >> 3, i = k & 7, t = n >> 4;
    int lane = g * 16 + (n & 15);
    return ((s * (N >> 4) + t) * 64 + lane) * 8 + i;
}

// The network's shape for hidden width WD (64 or 128): layer l maps K(l)
// inputs to N(l) outputs (64 padded features in, WD hidden, 48 out).
template <int WD>
struct GNet {
    static_assert(WD == 64 || WD == 128, "guide width must be 64 or 128");
    static constexpr int kWidth = WD;
    static CRH_GHD int K(int l) { return l == 0 ? kGIn : WD; }
    static CRH_GHD int N(int l) { return l < kGLayers - 1 ? WD : kGOut; }
    // Flat fp32 parameter vector: for each layer W (K x N, row-major) then b (N).
    static CRH_GHD int WOff(int l) {
        int off = 0;
        for (int i = 0; i < l; ++i) off += K(i) * N(i) + N(i);
        return off;
    }
    static CRH_GHD int BOff(int l) { return WOff(l) + K(l) * N(l); }
    // bf16 copies of the weights in MFMA B-operand order: W_l (K x N) and,
    // for the backward pass, W_l^T (N rounded up to 64 rows, padding rows
    // zero, x K); layer l takes K(l) x NP(l) elements in each copy.
    static CRH_GHD int NP(int l) { return (N(l) + 63) & ~63; }
    static CRH_GHD int QOff(int l) {
        int off = 0;
        for (int i = 0; i < l; ++i) off += K(i) * NP(i);
        return off;
    }
    // LDS row stride of a 16-row activation tile (+ 8 bf16: rows start 4
    // banks apart).
    static constexpr int kRow = WD + 8;
    static constexpr int kTile = 16 * kRow;
};
template <int WD>
constexpr int kGParamsOf = GNet<WD>::WOff(kGLayers);
template <int WD>
constexpr int kGQElemsOf = GNet<WD>::QOff(kGLayers);
static_assert(kGParamsOf<64> == 4 * (64 * 64 + 64) + (64 * 48 + 48), "19,760 parameters");
static_assert(kGParamsOf<128> == 64 * 128 + 128 + 3 * (128 * 128 + 128) + 128 * 48 + 48,
              "64,048 parameters");

// The default (64-wide) network's shape under its short names.
CRH_GHD int gLayerK(int l) { return GNet<64>::K(l); }
CRH_GHD int gLayerN(int l) { return GNet<64>::N(l); }
CRH_GHD int gWOff(int l) { return GNet<64>::WOff(l); }
CRH_GHD int gBOff(int l) { return GNet<64>::BOff(l); }
CRH_GHD int gQOff(int l) { return GNet<64>::QOff(l); }
constexpr int kGParams = kGParamsOf<64>;   // 19,760
constexpr int kGQElems = kGQElemsOf<64>;
// Parameters and operand elements of a network of width wd (64 or 128).
inline constexpr int gParamsOf(int wd) { return wd == 128 ? kGParamsOf<128> : kGParamsOf<64>; }
inline constexpr int gQElemsOf(int wd) { return wd == 128 ? kGQElemsOf<128> : kGQElemsOf<64>; }

// One training sample: shading point, normal, outgoing direction and the
// sampled incident direction; wi.w = target / sampling pdf.
struct alignas(16) GuideSample {
    float pos[4];
    float nrm[4];
    float wo[4];
    float wi[4];
};

}  // namespace crh
