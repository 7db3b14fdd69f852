// Device building blocks of the guiding network, shared by its kernels
// (guiding.hip) and the fused guided path kernel (guided_mk.hip), so that
// every one computes the same bits for a query: the input features, one
// layer of the MLP on v_mfma_f32_16x16x32_bf16 (BF16 operands, FP32
// accumulation) for a 16-row tile, and the vMF mixture head.
#pragma once
#include <hip/hip_runtime.h>

#include "guiding.h"

namespace crh {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kRow = GNet<64>::kRow;    // LDS row stride of a 64-wide network's tile
constexpr int kTile = GNet<64>::kTile;  // elements per activation tile

__device__ __forceinline__ __bf16 tobf(float v) { return static_cast<__bf16>(v); }
__device__ __forceinline__ float frombf(__bf16 v) { return static_cast<float>(v); }

// The head's single-precision building blocks on the hardware's
// transcendental and reciprocal instructions (v_exp_f32, v_log_f32,
// v_rcp_f32, v_sqrt_f32, about 1 ulp each): a head value is within a few
// 1e-7 (relative) of the correctly rounded one, far below what the bf16
// network resolves, at a fraction of the library functions' cost.
__device__ __forceinline__ float gExp(float x) {
    return __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
}
__device__ __forceinline__ float gLog(float x) {
    return __builtin_amdgcn_logf(x) * 0.693147180559945309f;
}
__device__ __forceinline__ float gRcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float gSqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
// log(1 + e), e >= 0, and exp(x) - 1, x <= 0, with Kahan's correction for
// the rounding of 1 + e (of exp(x)).
__device__ __forceinline__ float gLog1p(float e) {
    const float y = 1.f + e;
    return y == 1.f ? e : gLog(y) * (e * gRcp(y - 1.f));
}
__device__ __forceinline__ float gExpm1(float x) {
    const float u = gExp(x);
    if (u == 1.f) return x;
    const float y = u - 1.f;
    return y == -1.f ? -1.f : y * (x * gRcp(gLog(u)));
}

__device__ __forceinline__ float softplusf(float x) {
    return x > 20.f ? x : gLog1p(gExp(fminf(x, 20.f)));
}

// Input features of sample m of the tile into row m (lanes cover 16
// features each; rows ROW elements apart).  valid = false writes a zero row.
// Feature 10a + 2l (+ 1) is sin (cos) of 2^l pi p_a, p_a the position's
// offset along axis a over the box extent, l < 5; 32-34 the normal, 35-37
// the outgoing direction, the rest zero.  A lane of chunk 0 holds features
// 0-15 (axes 0 and 1), of chunk 1 features 16-31 (axes 1 and 2): it takes
// sincos of pi p for its two axes once and the higher levels by angle
// doubling (sin 2x = 2 sin x cos x, cos 2x = (cos x - sin x)(cos x + sin x)),
// within a few single-precision ulps of the direct values, far below the
// bf16 rounding of the features.
template <int ROW = kRow>
__device__ __forceinline__ void featureRow(__bf16* H, int m, int chunk, bool valid, const float* pos,
                           const float* nrm, const float* wo, const float* bmin,
                           const float* bext) {
    // chunk varies per lane: the three-element arrays are read with selects,
    // not indexed, so they stay in registers
    auto pick = [](const float* v, int i) { return i == 0 ? v[0] : (i == 1 ? v[1] : v[2]); };
    float sn[2][5], cs[2][5];  // the lane's axes chunk and chunk + 1 (chunks 2, 3: unused)
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int a = min(chunk + q, 2);
        const float ea = pick(bext, a);
        const float e = ea > 0.f ? ea : 1.f;
        const float p = (pick(pos, a) - pick(bmin, a)) / e;
        sincosf(3.14159265358979323846f * p, &sn[q][0], &cs[q][0]);
#pragma unroll
        for (int l = 1; l < 5; ++l) {
            sn[q][l] = 2.f * sn[q][l - 1] * cs[q][l - 1];
            cs[q][l] = (cs[q][l - 1] - sn[q][l - 1]) * (cs[q][l - 1] + sn[q][l - 1]);
        }
    }
    // feature f (< 30) of a lane whose first axis is a0: axis f / 10
    auto freq = [&](int f, int a0) {
        const int q = f / 10 - a0, l = (f % 10) >> 1;
        return (f & 1) ? cs[q][l] : sn[q][l];
    };
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        // the candidates of chunks 0 .. 3 for this j (indices known at compile time)
        const float v0 = freq(j, 0);
        const float v1 = 16 + j < 30 ? freq(16 + j, 1) : 0.f;
        const float v2 = j < 3 ? pick(nrm, j) : (j < 6 ? pick(wo, j - 3) : 0.f);
        float v = chunk == 0 ? v0 : (chunk == 1 ? v1 : (chunk == 2 ? v2 : 0.f));
        H[m * ROW + chunk * 16 + j] = tobf(valid ? v : 0.f);
    }
}

// One layer of the width-WD network: out = act(in * W_l + b_l) for the
// wave's 16-row tile (rows GNet<WD>::kRow apart), K(l) / 32 MFMA steps per
// 16 output columns, in k order; wq: the forward operand copy (LDS or
// global memory).  For the last layer the fp32 result goes to outF
// ([16][48]).
template <int WD>
__device__ __forceinline__ void layerForward(int l, const __bf16* in, __bf16* out, float* outF,
                                             const __bf16* wq, const float* __restrict__ theta) {
    using G = GNet<WD>;
    const int lane = threadIdx.x & 63, g = lane >> 4, r16 = lane & 15;
    const int N = G::N(l), NT = N >> 4, KS = G::K(l) >> 5;
    const float* bias = theta + G::BOff(l);
    for (int t = 0; t < NT; ++t) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int s = 0; s < KS; ++s) {
            bf16x8 a = *reinterpret_cast<const bf16x8*>(in + r16 * G::kRow + 32 * s + 8 * g);
            bf16x8 b = *reinterpret_cast<const bf16x8*>(wq + G::QOff(l) +
                                                         ((s * NT + t) * 64 + lane) * 8);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
        const int col = 16 * t + r16;
        const float bc = bias[col];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = 4 * g + i;
            float v = acc[i] + bc;
            if (l < kGLayers - 1)
                out[row * G::kRow + col] = tobf(fmaxf(v, 0.f));
            else
                outF[row * kGOut + col] = v;
        }
    }
}

struct Head {
    float w[kGLobes], mu[kGLobes][3], k[kGLobes], kraw[kGLobes], mlen[kGLobes];
};

// The head lobe by lobe (every user computes a lobe's values with these):
// softmax weight (largest logit subtracted, the exponentials summed in lobe
// order), unit mean (length floored at 1e-6, the components scaled by its
// reciprocal) and kappa = softplus + 0.01.
__device__ __forceinline__ float headLogitMax(const float* o) {
    float amax = -INFINITY;
#pragma unroll
    for (int j = 0; j < kGLobes; ++j) amax = fmaxf(amax, o[5 * j]);
    return amax;
}
__device__ __forceinline__ float headExpSum(const float* o, float amax) {
    float sum = 0.f;
#pragma unroll
    for (int j = 0; j < kGLobes; ++j) sum += gExp(o[5 * j] - amax);
    return sum;
}
__device__ __forceinline__ float headWeight(const float* o, int j, float amax, float sum) {
    return gExp(o[5 * j] - amax) * gRcp(sum);
}
__device__ __forceinline__ float headMeanLength(const float* o, int j) {
    const float mx = o[5 * j + 1], my = o[5 * j + 2], mz = o[5 * j + 3];
    return fmaxf(gSqrt(mx * mx + my * my + mz * mz), 1e-6f);
}
__device__ __forceinline__ float headKappa(const float* o, int j) {
    return softplusf(o[5 * j + 4]) + kGKappaMin;
}

__device__ __forceinline__ void computeHead(const float* o, Head& h) {
    const float amax = headLogitMax(o), sum = headExpSum(o, amax);
#pragma unroll
    for (int j = 0; j < kGLobes; ++j) {
        h.w[j] = headWeight(o, j, amax, sum);
        const float len = headMeanLength(o, j), il = gRcp(len);
        h.mlen[j] = len;
        h.mu[j][0] = o[5 * j + 1] * il;
        h.mu[j][1] = o[5 * j + 2] * il;
        h.mu[j][2] = o[5 * j + 3] * il;
        h.kraw[j] = o[5 * j + 4];
        h.k[j] = headKappa(o, j);
    }
}

__device__ __forceinline__ float vmfPdf(float k, float c) {
    // k / (2 pi (1 - e^{-2k})) * exp(k (c - 1))
    return k * gRcp(2.f * 3.14159265358979323846f * (-gExpm1(-2.f * k))) * gExp(k * (c - 1.f));
}

}  // namespace crh
