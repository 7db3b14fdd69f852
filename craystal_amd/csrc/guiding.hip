// Path-guiding network kernels for gfx950: inference and training of the
// MLP (hidden width 64 or 128, template WD) with v_mfma_f32_16x16x32_bf16
// (BF16 operands, FP32 accumulation), the von Mises-Fisher mixture head and
// Adam.
//
// Tiles: one wave computes 16 samples; a block of 4 waves moves through the
// batch in lockstep (64 samples per iteration) so block barriers stay in
// uniform control flow.  Activations are row-major [16][WD + 8] bf16 tiles
// in LDS (8 bf16 of padding: row starts 4 banks apart).  The 64-wide
// network's forward weights (40 KB) are staged once per block in LDS,
// already in MFMA B-operand order, so each operand is one 16-byte LDS read
// per lane; the 128-wide network's (128 KB) are read from global memory
// (L2) in the same order.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "guide_mlp.h"
#include "guiding.h"
#include "guiding_kernels.h"

namespace crh {

constexpr int kWavesPerBlock = 4;

namespace {

// Whether the width-WD network's forward weights are staged in LDS.
template <int WD>
constexpr bool kStageW = WD == 64;
template <int WD>
constexpr int kStagedElems = kStageW<WD> ? kGQElemsOf<WD> : 0;

// Stage the forward weights in LDS (64-wide network), or use them in
// global memory; returns where the layers read them.
template <int WD>
__device__ const __bf16* stageWeights(const __bf16* __restrict__ wq, __bf16* ldsW) {
    if constexpr (!kStageW<WD>) return wq;
    const uint4* src = reinterpret_cast<const uint4*>(wq);
    uint4* dst = reinterpret_cast<uint4*>(ldsW);
    for (int i = threadIdx.x; i < kGQElemsOf<WD> / 8; i += blockDim.x) dst[i] = src[i];
    return ldsW;
}

}  // namespace

// ---------------------------------------------------------------------------
// Inference: per queued path, mixture parameters
//   out[p * 40 + ...] = w[8], mu[8][3], kappa[8]
// ---------------------------------------------------------------------------
template <int WD>
__global__ __launch_bounds__(256) void kGuideInfer(GuideView gv, const uint32_t* __restrict__ queue,
                                                   const uint32_t* __restrict__ count,
                                                   uint32_t countDirect,
                                                   const float4* __restrict__ pos,
                                                   const float4* __restrict__ nrm,
                                                   const float4* __restrict__ wo,
                                                   float* __restrict__ params,
                                                   float* __restrict__ rawOut) {
    using G = GNet<WD>;
    extern __shared__ __align__(16) unsigned char smem[];
    __bf16* ldsW = reinterpret_cast<__bf16*>(smem);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __bf16* H0 = ldsW + kStagedElems<WD> + wave * (2 * G::kTile);
    __bf16* H1 = H0 + G::kTile;
    float* outF = reinterpret_cast<float*>(ldsW + kStagedElems<WD> +
                                           kWavesPerBlock * 2 * G::kTile) +
                  wave * 16 * kGOut;
    const __bf16* W = stageWeights<WD>(reinterpret_cast<const __bf16*>(gv.wq), ldsW);
    const uint32_t n = count ? *count : countDirect;
    for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
        __syncthreads();
        const uint32_t tile0 = base + wave * 16;
        {
            int m = lane & 15, chunk = lane >> 4;
            uint32_t idx = tile0 + m;
            bool valid = idx < n;
            uint32_t p = valid ? (queue ? queue[idx] : idx) : 0u;
            float4 P = valid ? pos[p] : make_float4(0, 0, 0, 0);
            float4 Nn = valid ? nrm[p] : make_float4(0, 0, 0, 0);
            float4 W = valid ? wo[p] : make_float4(0, 0, 0, 0);
            const float pa[3] = {P.x, P.y, P.z}, na[3] = {Nn.x, Nn.y, Nn.z},
                        wa[3] = {W.x, W.y, W.z};
            featureRow<G::kRow>(H0, m, chunk, valid, pa, na, wa, gv.bmin, gv.bext);
        }
        __syncthreads();
        __bf16* in = H0;
        __bf16* out = H1;
        for (int l = 0; l < kGLayers; ++l) {
            layerForward<WD>(l, in, out, outF, W, gv.theta);
            __syncthreads();
            __bf16* t = in;
            in = out;
            out = t;
        }
        if (lane < 16) {
            uint32_t idx = tile0 + lane;
            if (idx < n) {
                uint32_t p = queue ? queue[idx] : idx;
                const float* o = outF + lane * kGOut;
                if (rawOut)
                    for (int c = 0; c < kGOut; ++c) rawOut[size_t(p) * kGOut + c] = o[c];
                Head h;
                computeHead(o, h);
                float* dst = params + size_t(p) * 40;
                for (int j = 0; j < kGLobes; ++j) {
                    dst[j] = h.w[j];
                    dst[8 + 3 * j] = h.mu[j][0];
                    dst[8 + 3 * j + 1] = h.mu[j][1];
                    dst[8 + 3 * j + 2] = h.mu[j][2];
                    dst[32 + j] = h.k[j];
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Training, pass 1: forward, KL-loss gradient at the outputs, backward
// through the layers; saves every layer's input H_l and output gradient
// delta_l (bf16 rows of WD; H_0 and delta_4 use their first 64 and 48
// columns) for the weight-gradient GEMM.
// ---------------------------------------------------------------------------
template <int WD>
__global__ __launch_bounds__(256) void kGuideFwdBwd(GuideView gv,
                                                    const GuideSample* __restrict__ samples,
                                                    uint32_t n, float invN,
                                                    __bf16* __restrict__ Hsave,
                                                    __bf16* __restrict__ Dsave,
                                                    float* __restrict__ lossSum) {
    using G = GNet<WD>;
    constexpr int kT = G::kTile, kR = G::kRow;
    extern __shared__ __align__(16) unsigned char smem[];
    __bf16* ldsW = reinterpret_cast<__bf16*>(smem);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4,
              r16 = lane & 15;
    // per wave: H_0..H_4 (5 tiles) + two delta tiles
    __bf16* Hw = ldsW + kStagedElems<WD> + wave * (7 * kT);
    __bf16* Da = Hw + 5 * kT;
    __bf16* Db = Da + kT;
    float* outF = reinterpret_cast<float*>(ldsW + kStagedElems<WD> + kWavesPerBlock * 7 * kT) +
                  wave * 16 * kGOut;
    const __bf16* W = stageWeights<WD>(reinterpret_cast<const __bf16*>(gv.wq), ldsW);
    const size_t padN = (size_t(n) + 63) & ~size_t(63);
    for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
        __syncthreads();
        const uint32_t tile0 = base + wave * 16;
        {
            int m = lane & 15, chunk = lane >> 4;
            uint32_t idx = tile0 + m;
            bool valid = idx < n;
            const GuideSample& smp = samples[valid ? idx : 0];
            featureRow<kR>(Hw, m, chunk, valid, smp.pos, smp.nrm, smp.wo, gv.bmin, gv.bext);
        }
        __syncthreads();
        for (int l = 0; l < kGLayers; ++l) {
            layerForward<WD>(l, Hw + l * kT, l + 1 < kGLayers ? Hw + (l + 1) * kT : nullptr,
                             outF, W, gv.theta);
            __syncthreads();
        }
        // Output gradient (oracle/guiding.py loss_and_grad_out), lanes 0..15.
        for (int c = lane; c < 16 * kR; c += 64) Da[c] = tobf(0.f);
        __syncthreads();
        if (lane < 16) {
            uint32_t idx = tile0 + lane;
            if (idx < n) {
                const GuideSample& smp = samples[idx];
                const float* o = outF + lane * kGOut;
                Head h;
                computeHead(o, h);
                float d[3] = {smp.wi[0], smp.wi[1], smp.wi[2]};
                float v[kGLobes], q = 0.f, cmu[kGLobes];
                for (int j = 0; j < kGLobes; ++j) {
                    cmu[j] = h.mu[j][0] * d[0] + h.mu[j][1] * d[1] + h.mu[j][2] * d[2];
                    v[j] = vmfPdf(h.k[j], cmu[j]);
                    q += h.w[j] * v[j];
                }
                q = fmaxf(q, 1e-30f);
                float wt = smp.wi[3];
                atomicAdd(lossSum, -wt * logf(q) * invN);
                float gsc = -wt * invN;
                for (int j = 0; j < kGLobes; ++j) {
                    float resp = h.w[j] * v[j] / q;
                    float ga = gsc * (resp - h.w[j]);
                    float e = expf(-2.f * h.k[j]);
                    float dlv = 1.f / h.k[j] - 2.f * e / (1.f - e) + (cmu[j] - 1.f);
                    float sig = 1.f / (1.f + expf(-h.kraw[j]));
                    float gk = gsc * resp * dlv * sig;
                    float s = gsc * resp * h.k[j] / h.mlen[j];
                    __bf16* row = Da + lane * kR;
                    row[5 * j] = tobf(ga);
                    for (int c = 0; c < 3; ++c)
                        row[5 * j + 1 + c] = tobf(s * (d[c] - cmu[j] * h.mu[j][c]));
                    row[5 * j + 4] = tobf(gk);
                }
            }
        }
        __syncthreads();
        // Save H_l rows (K(l) columns) and delta_4 (64 columns, 48 used);
        // then backward: delta_{l-1} = (delta_l W_l^T) masked by H_l > 0.
        for (int c = lane; c < 16 * WD; c += 64) {
            int m = c / WD, f = c % WD;
            uint32_t idx = tile0 + m;
            if (idx < padN) {
                for (int l = 0; l < kGLayers; ++l)
                    if (f < G::K(l))
                        Hsave[(size_t(l) * padN + idx) * WD + f] = Hw[l * kT + m * kR + f];
                if (f < G::NP(kGLayers - 1))
                    Dsave[(size_t(4) * padN + idx) * WD + f] = Da[m * kR + f];
            }
        }
        __bf16* dIn = Da;
        __bf16* dOut = Db;
        for (int l = kGLayers - 1; l >= 1; --l) {
            const __bf16* H = Hw + l * kT;
            const int KT = G::K(l) >> 4, NS = G::NP(l) >> 5;
            for (int t = 0; t < KT; ++t) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                for (int s = 0; s < NS; ++s) {
                    bf16x8 a = *reinterpret_cast<const bf16x8*>(dIn + r16 * kR + 32 * s + 8 * g);
                    bf16x8 b = *reinterpret_cast<const bf16x8*>(
                        reinterpret_cast<const __bf16*>(gv.wqT) + G::QOff(l) +
                        ((s * KT + t) * 64 + lane) * 8);
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
                }
                const int col = 16 * t + r16;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int row = 4 * g + i;
                    float hv = frombf(H[row * kR + col]);
                    __bf16 dv = tobf(hv > 0.f ? acc[i] : 0.f);
                    dOut[row * kR + col] = dv;
                    uint32_t idx = tile0 + row;
                    if (idx < padN) Dsave[(size_t(l - 1) * padN + idx) * WD + col] = dv;
                }
            }
            __syncthreads();
            __bf16* t2 = dIn;
            dIn = dOut;
            dOut = t2;
        }
    }
}

// ---------------------------------------------------------------------------
// Training, pass 2: dW_l = H_l^T delta_l over the batch (K = samples, 32 per
// MFMA step) and db_l = sum delta_l, accumulated into grad with atomics.
// grid.x = (WD / 16)^2 output tiles per layer x 5 layers (tiles outside a
// layer's K x N leave at once), grid.y = batch chunks.
// ---------------------------------------------------------------------------
template <int WD>
__global__ __launch_bounds__(64) void kGuideWeightGrad(GuideView gv, const __bf16* __restrict__ Hsave,
                                                       const __bf16* __restrict__ Dsave, uint32_t n,
                                                       uint32_t chunk) {
    using G = GNet<WD>;
    constexpr int TS = WD / 16;  // tiles per side
    const int lane = threadIdx.x & 63, g = lane >> 4, r16 = lane & 15;
    const int l = blockIdx.x / (TS * TS), tile = blockIdx.x % (TS * TS), a = tile / TS,
              b = tile % TS;
    const int N = G::N(l);
    if (16 * a >= G::K(l) || 16 * b >= N) return;
    const size_t padN = (size_t(n) + 63) & ~size_t(63);
    const uint32_t s0 = blockIdx.y * chunk, s1 = min(n, s0 + chunk);
    const __bf16* H = Hsave + size_t(l) * padN * WD;
    const __bf16* D = Dsave + size_t(l) * padN * WD;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;
    for (uint32_t s = s0; s < s1; s += 32) {
        bf16x8 av, bv;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            uint32_t smp = s + 8 * g + i;
            bool ok = smp < s1;
            av[i] = ok ? H[size_t(smp) * WD + 16 * a + r16] : tobf(0.f);
            bv[i] = ok ? D[size_t(smp) * WD + 16 * b + r16] : tobf(0.f);
            if (a == 0) bsum += frombf(bv[i]);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    }
    float* grad = gv.grad;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int row = 16 * a + 4 * g + i, col = 16 * b + r16;
        atomicAdd(&grad[G::WOff(l) + row * N + col], acc[i]);
    }
    if (a == 0) {
        // lanes with the same column (r16) hold partial sums for g = 0..3
        bsum += __shfl_xor(bsum, 16, 64);
        bsum += __shfl_xor(bsum, 32, 64);
        if (g == 0) atomicAdd(&grad[G::BOff(l) + 16 * b + r16], bsum);
    }
}

// ---------------------------------------------------------------------------
// Adam on the fp32 master parameters; refreshes both bf16 operand copies
// and clears the gradient for the next step.
// ---------------------------------------------------------------------------
template <int WD>
__global__ __launch_bounds__(256) void kGuideAdam(GuideView gv, float lr, float b1, float b2,
                                                  float eps, float c1, float c2) {
    using G = GNet<WD>;
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kGParamsOf<WD>) return;
    float gr = gv.grad[i];
    float m = b1 * gv.m[i] + (1.f - b1) * gr;
    float v = b2 * gv.v[i] + (1.f - b2) * gr * gr;
    gv.m[i] = m;
    gv.v[i] = v;
    float th = gv.theta[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    gv.theta[i] = th;
    gv.grad[i] = 0.f;
    // weight element?  find its layer
    for (int l = 0; l < kGLayers; ++l) {
        int w0 = G::WOff(l), K = G::K(l), N = G::N(l);
        if (i >= w0 && i < w0 + K * N) {
            int k = (i - w0) / N, c = (i - w0) % N;
            reinterpret_cast<__bf16*>(gv.wq)[G::QOff(l) + gSwz(k, c, N)] = tobf(th);
            reinterpret_cast<__bf16*>(gv.wqT)[G::QOff(l) + gSwz(c, k, K)] = tobf(th);
        }
    }
}

// Rebuild the bf16 operand copies from theta (after a host upload): one
// thread per element of the copies; W^T's padding rows stay zero.
template <int WD>
__global__ __launch_bounds__(256) void kGuideQuantize(GuideView gv) {
    using G = GNet<WD>;
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kGQElemsOf<WD>) return;
    reinterpret_cast<__bf16*>(gv.wqT)[i] = tobf(0.f);
    for (int l = 0; l < kGLayers; ++l) {
        int w0 = G::WOff(l), N = G::N(l);
        int j = i - G::QOff(l);
        if (j >= 0 && j < G::K(l) * N) {
            int k = j / N, c = j % N;
            float th = gv.theta[w0 + j];
            reinterpret_cast<__bf16*>(gv.wq)[G::QOff(l) + gSwz(k, c, N)] = tobf(th);
        }
    }
}
template <int WD>
__global__ __launch_bounds__(256) void kGuideQuantizeT(GuideView gv) {
    using G = GNet<WD>;
    int i = blockIdx.x * 256 + threadIdx.x;
    for (int l = 0; l < kGLayers; ++l) {
        int w0 = G::WOff(l), K = G::K(l), N = G::N(l);
        int j = i - G::QOff(l);
        if (j >= 0 && j < K * N) {
            int k = j / N, c = j % N;
            reinterpret_cast<__bf16*>(gv.wqT)[G::QOff(l) + gSwz(c, k, K)] =
                tobf(gv.theta[w0 + j]);
        }
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
namespace {

template <int WD>
size_t inferLds() {
    return size_t(kStagedElems<WD>) * 2 + size_t(kWavesPerBlock) * 2 * GNet<WD>::kTile * 2 +
           size_t(kWavesPerBlock) * 16 * kGOut * 4;
}
template <int WD>
size_t trainLds() {
    return size_t(kStagedElems<WD>) * 2 + size_t(kWavesPerBlock) * 7 * GNet<WD>::kTile * 2 +
           size_t(kWavesPerBlock) * 16 * kGOut * 4;
}

// f(integral_constant<int, WD>) for the view's width.
template <class F>
void withWidth(uint32_t width, F&& f) {
    if (width == 128) f(std::integral_constant<int, 128>{});
    else f(std::integral_constant<int, 64>{});
}

}  // namespace

hipError_t guideSetLdsLimits() {
    hipError_t e = hipSuccess;
    auto set = [&](const void* k, size_t bytes) {
        if (e == hipSuccess)
            e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes));
    };
    set(reinterpret_cast<const void*>(&kGuideInfer<64>), inferLds<64>());
    set(reinterpret_cast<const void*>(&kGuideInfer<128>), inferLds<128>());
    set(reinterpret_cast<const void*>(&kGuideFwdBwd<64>), trainLds<64>());
    set(reinterpret_cast<const void*>(&kGuideFwdBwd<128>), trainLds<128>());
    return e;
}

hipError_t launchGuideInfer(const GuideView& gv, const uint32_t* queue, const uint32_t* count,
                            uint32_t countDirect, const float4* pos, const float4* nrm,
                            const float4* wo, float* params, float* rawOut, uint32_t grid,
                            hipStream_t st) {
    withWidth(gv.width, [&](auto w) {
        constexpr int WD = decltype(w)::value;
        hipLaunchKernelGGL(kGuideInfer<WD>, dim3(grid), dim3(256), inferLds<WD>(), st, gv, queue,
                           count, countDirect, pos, nrm, wo, params, rawOut);
    });
    return hipGetLastError();
}

hipError_t launchGuideQuantize(const GuideView& gv, hipStream_t st) {
    withWidth(gv.width, [&](auto w) {
        constexpr int WD = decltype(w)::value;
        const dim3 grid((kGQElemsOf<WD> + 255) / 256);
        hipLaunchKernelGGL(kGuideQuantize<WD>, grid, dim3(256), 0, st, gv);
        hipLaunchKernelGGL(kGuideQuantizeT<WD>, grid, dim3(256), 0, st, gv);
    });
    return hipGetLastError();
}

hipError_t launchGuideTrainStep(const GuideView& gv, const GuideSample* samples, uint32_t n,
                                uint16_t* Hsave16, uint16_t* Dsave16, float* lossSum, float lr,
                                uint32_t step, bool applyAdam, hipStream_t st) {
    __bf16* Hsave = reinterpret_cast<__bf16*>(Hsave16);
    __bf16* Dsave = reinterpret_cast<__bf16*>(Dsave16);
    if (n == 0) return hipSuccess;
    uint32_t blocks = (n + 63) / 64;
    if (blocks > 512) blocks = 512;
    withWidth(gv.width, [&](auto w) {
        constexpr int WD = decltype(w)::value;
        hipLaunchKernelGGL(kGuideFwdBwd<WD>, dim3(blocks), dim3(256), trainLds<WD>(), st, gv,
                           samples, n, 1.0f / float(n), Hsave, Dsave, lossSum);
        const uint32_t chunk = 2048;
        dim3 wg((WD / 16) * (WD / 16) * kGLayers, (n + chunk - 1) / chunk);
        hipLaunchKernelGGL(kGuideWeightGrad<WD>, wg, dim3(64), 0, st, gv, Hsave, Dsave, n, chunk);
        if (applyAdam) {
            const float b1 = 0.9f, b2 = 0.999f;
            float c1 = 1.f - powf(b1, float(step)), c2 = 1.f - powf(b2, float(step));
            hipLaunchKernelGGL(kGuideAdam<WD>, dim3((kGParamsOf<WD> + 255) / 256), dim3(256), 0,
                               st, gv, lr, b1, b2, 1e-8f, c1, c2);
        }
    });
    return hipGetLastError();
}

}  // namespace crh
