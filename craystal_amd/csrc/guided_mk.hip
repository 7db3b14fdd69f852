// Fused guided path tracing for gfx950: the path-regenerating megakernel
// (kRenderRegen) with the guide's inference inside.  Every iteration each
// live lane runs one bounce of its path: the closest hit, the emitter or
// environment term and the light sample with MIS, as pathBounce does.  Then
// the wave evaluates the guide for its lanes whose vertex is guided, on
// v_mfma_f32_16x16x32_bf16: the guided lanes' queries are packed into as
// few 16-row tiles as they fill, each row's features built in the wave's
// LDS tile, five layers computed in place (A operands loaded before the
// tile is overwritten), the weights read in MFMA operand order from L2.
// Each guided lane finds its row's 40 outputs in a per-wave scratch buffer,
// computes the mixture head, and makes the continuation: one-sample MIS of BSDF sampling and the
// mixture (guidedSampleEval), or the plain BSDF sample at other vertices,
// then the kill and roulette.  A row's outputs do not depend on the other
// rows of its tile, and every step and draw is the wavefront guided path's
// (kWfShadeCont), so for the same guide the films are bit-identical to it.
//
// Training records are the wavefront's: one per guided vertex with a
// nonzero pdf (shading point, q_mix, normal, wo, wi, f, throughput), the
// radiance at the start of the path's next bounce (the snapshot), and the
// final radiance read from L at conversion.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "guide_device.h"
#include "guide_mlp.h"
#include "guided_mk.h"
#include "kernels.h"
#include "path.h"
#include "wave.h"
#include "wavefront.h"
#include "batch_device.h"

#ifndef CRH_GMK_PART
#error "compile with -DCRH_GMK_PART=<0|1|2> (kernels of one stack mode) or 3 (the rest)"
#endif

namespace crh {

namespace {

// A wave's 16-row bf16 activation tile of the width-WD network.
template <int WD>
constexpr uint32_t kTileBytesOf = GNet<WD>::kTile * 2;

// Orders the LDS accesses of the lanes of a wave across this point (LDS
// only: the wave's global loads stay in flight).
__device__ __forceinline__ void ldsSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Orders LDS and global accesses of the lanes of a wave (and of its
// workgroup's shared L1) across this point.
__device__ __forceinline__ void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Buffer resource of `bytes` bytes at p (raw buffer, gfx9 word 3): buffer
// loads keep the base in SGPRs and take a lane offset plus a scalar offset,
// so the weights' addresses cost no vector registers.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t bufferOf(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, int(bytes), 0x00020000);
}

// Layer L of the width-WD MLP on the wave's 16-row tile H, in place:
// hidden layers write bf16 ReLU outputs back into H after every lane has
// read its A operands, the output layer writes row r's first kGuideRaw
// outputs to floats outOff / 4 + r * kGuideRaw .. of the buffer `out`.  The
// 64-wide layers load all their weights (in MFMA operand order) and biases
// in one batch before the A operands; the 128-wide ones (4 x 8 operand
// blocks) load each 16-column block's as they go.  The arithmetic, and its
// order, is layerForward's.
template <int WD, int L>
__device__ __forceinline__ void layerInPlace(__bf16* H, __amdgpu_buffer_rsrc_t wq,
                                             __amdgpu_buffer_rsrc_t theta,
                                             __amdgpu_buffer_rsrc_t out, uint32_t outOff) {
    using G = GNet<WD>;
    constexpr int NT = G::N(L) >> 4, KS = G::K(L) >> 5, R = G::kRow;
    constexpr bool kBatch = WD == 64;
    const int lane = threadIdx.x & 63, g = lane >> 4, r16 = lane & 15;
    auto ldB = [&](int s, int t) {
        return __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(
                                              wq, lane * 16, (G::QOff(L) + (s * NT + t) * 64 * 8) * 2, 0));
    };
    auto ldBias = [&](int t) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                             theta, r16 * 4, (G::BOff(L) + 16 * t) * 4, 0));
    };
    bf16x8 bAll[kBatch ? KS : 1][kBatch ? NT : 1];
    float bcAll[kBatch ? NT : 1];
    if constexpr (kBatch) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
#pragma unroll
            for (int s = 0; s < KS; ++s) bAll[s][t] = ldB(s, t);
            bcAll[t] = ldBias(t);
        }
    }
    bf16x8 a[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s)
        a[s] = *reinterpret_cast<const bf16x8*>(H + r16 * R + 32 * s + 8 * g);
    if (L < kGLayers - 1) ldsSync();
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        bf16x8 b[KS];
        float bc;
        if constexpr (kBatch) {
#pragma unroll
            for (int s = 0; s < KS; ++s) b[s] = bAll[s][t];
            bc = bcAll[t];
        } else {
#pragma unroll
            for (int s = 0; s < KS; ++s) b[s] = ldB(s, t);
            bc = ldBias(t);
        }
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < KS; ++s)
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[s], b[s], acc, 0, 0, 0);
        const int col = 16 * t + r16;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = 4 * g + i;
            const float v = acc[i] + bc;
            if (L < kGLayers - 1)
                H[row * R + col] = tobf(fmaxf(v, 0.f));
            else if (col < int(kGuideRaw))
                __builtin_amdgcn_raw_buffer_store_b32(
                    __builtin_bit_cast(uint32_t, v), out, (4 * g * int(kGuideRaw) + r16) * 4,
                    outOff + uint32_t(i * int(kGuideRaw) + 16 * t) * 4u, 0);
        }
    }
    if (L < kGLayers - 1) ldsSync();
}

template <int WD, int L>
__device__ __forceinline__ void layersFrom(__bf16* H, __amdgpu_buffer_rsrc_t wq,
                                           __amdgpu_buffer_rsrc_t theta,
                                           __amdgpu_buffer_rsrc_t out, uint32_t outOff) {
    layerInPlace<WD, L>(H, wq, theta, out, outOff);
    if constexpr (L + 1 < kGLayers) layersFrom<WD, L + 1>(H, wq, theta, out, outOff);
}

// Position of set bit k (0-based, in ascending order) of m, for k below
// popcount(m): the largest p with fewer than k + 1 set bits below p.
__device__ __forceinline__ int kthSetBit(uint64_t m, uint32_t k) {
    int pos = 0;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1) {
        const uint64_t below = m & ((uint64_t(1) << (pos + step)) - 1u);
        if (uint32_t(__popcll(below)) <= k) pos += step;
    }
    return pos;
}

// The guide for the wave's lanes with `guided` set, packed into as few
// 16-row tiles as they fill: the guided lane of rank r (r-th guided lane
// of the wave in lane order) is row r % 16 of tile r / 16 and gets its
// kGuideRaw outputs at raw[r * kGuideRaw ..] (the wave's 64 x kGuideRaw
// floats); returns this lane's rank.  A row's outputs do not depend on the
// other rows of its tile, so the packing changes no value.  Wave-uniform
// control flow.
template <int WD>
__device__ __forceinline__ uint32_t guideWave(const GuideView& gv, bool guided, F3 pos, F3 nrm,
                                              F3 wo, unsigned char* tile,
                                              float* __restrict__ raw) {
    const uint64_t gm = __ballot(guided);
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(gm >> 32),
                                                    __builtin_amdgcn_mbcnt_lo(uint32_t(gm), 0u));
    if (gm == 0) return rank;
    const int lane = threadIdx.x & 63, m = lane & 15, chunk = lane >> 4;
    const int n = __popcll(gm);
    const int srcOfRow = kthSetBit(gm, uint32_t(lane));  // the lane of rank `lane`
    __bf16* H = reinterpret_cast<__bf16*>(tile);
    const __amdgpu_buffer_rsrc_t wq = bufferOf(gv.wq, uint32_t(kGQElemsOf<WD>) * 2u);
    const __amdgpu_buffer_rsrc_t theta = bufferOf(gv.theta, uint32_t(kGParamsOf<WD>) * 4u);
    const __amdgpu_buffer_rsrc_t out = bufferOf(raw, 64u * kGuideRaw * 4u);
    for (int t = 0; 16 * t < n; ++t) {
        const int row = 16 * t + m;
        const bool valid = row < n;
        const int src = __shfl(srcOfRow, row, 64);
        const float pa[3] = {__shfl(pos.x, src, 64), __shfl(pos.y, src, 64), __shfl(pos.z, src, 64)};
        const float na[3] = {__shfl(nrm.x, src, 64), __shfl(nrm.y, src, 64), __shfl(nrm.z, src, 64)};
        const float wa[3] = {__shfl(wo.x, src, 64), __shfl(wo.y, src, 64), __shfl(wo.z, src, 64)};
        featureRow<GNet<WD>::kRow>(H, m, chunk, valid, pa, na, wa, gv.bmin, gv.bext);
        ldsSync();
        layersFrom<WD, 0>(H, wq, theta, out, uint32_t(16 * t) * kGuideRaw * 4u);
    }
    waveSync();  // the outputs, written by other lanes, are read next
    return rank;
}

// MULTI: several camera views of the scene in one launch (GuidedViews):
// path p is tile (p % perSlot) >> 8 of all views' tiles, view after view,
// of sample slot p / perSlot, and writes slot * totalPix + pixPrefix[view] +
// its pixel of L and P; otherwise one camera and slot * npix + pixel.
template <bool MIS, int SM, int W, bool MULTI, int WD>
__global__ __launch_bounds__(kRenderBlock) __attribute__((amdgpu_waves_per_eu(W, 8))) void
kRenderGuided(SceneView s, CameraView cam, IntegratorView cfg, uint32_t si, uint32_t tilesX,
              uint32_t tiles, uint32_t total, uint32_t* __restrict__ next, float4* __restrict__ L,
              float2* __restrict__ P, unsigned long long* __restrict__ counters, GuideView gv,
              float alpha, uint32_t gmask, uint32_t guideDepth, float* __restrict__ scratch,
              uint32_t tileOff, float4* __restrict__ rec, uint32_t* __restrict__ recCounter,
              uint32_t recCap, GuidedViews vs) {
    extern __shared__ uint2 lds[];
    auto stk = makeStack<SM>(lds);
    PathCtx<false, typename StackOf<SM>::T> c{s, stk, 0, {0, 0}};
    const uint32_t waveInBlock = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    unsigned char* tile =
        reinterpret_cast<unsigned char*>(lds) + tileOff + waveInBlock * kTileBytesOf<WD>;
    float* raw = scratch + size_t(blockIdx.x * (kRenderBlock / 64u) +
                                  __builtin_amdgcn_readfirstlane(waveInBlock)) *
                               64u * kGuideRaw;
    const size_t npix = size_t(cam.width) * cam.height;
    const uint32_t perSlot = tiles * 256u;
    const Lanes w = rgbLanes();
    bool live = false;
    WaveDispenser disp;
    const uint32_t chunk = dispenserChunk(total);
    uint32_t depth = 0, recId = kNoRecord;
    size_t out = 0;
    Sampler smp(0u, 0u, 0u);
    Ray ray{};
    Spec rad = sp(0.f), beta = sp(1.f);
    F2 pix{0.f, 0.f};
    auto finish = [&]() {
        L[out] = make_float4(rad.r, rad.g, rad.b, 0.f);
        P[out] = make_float2(pix.x, pix.y);
        live = false;
    };
    while (true) {
        {
            const uint32_t p = disp.take(next, total, chunk, !live);
            if (p < total) {
                const uint32_t slot = p / perSlot, q = p - slot * perSlot;
                uint32_t x, y, view = 0u;
                const CameraView* cv = &cam;
                if constexpr (MULTI) {
                    view = findScene(vs.tilePrefix, vs.n, q >> 8);
                    cv = &vs.cams[view];
                    tileCoordOf((q >> 8) - vs.tilePrefix[view], q & 255u,
                                (cv->width + 15u) / 16u, x, y);
                } else {
                    tileCoordOf(q >> 8, q & 255u, tilesX, x, y);
                }
                if (x < cv->width && y < cv->height) {
                    smp = Sampler(x, y, si + slot);
                    F2 j = smp.next2D();
                    pix = F2{float(x) + j.x, float(y) + j.y};
                    ray = cameraRay(*cv, pix);
                    rad = sp(0.f);
                    beta = sp(1.f);
                    depth = 0;
                    recId = kNoRecord;
                    out = MULTI ? size_t(slot) * vs.totalPix + vs.pixPrefix[view] +
                                      size_t(y) * cv->width + x
                                : slot * npix + size_t(y) * cam.width + x;
                    live = true;
                    if (cfg.maxDepth == 0u) finish();
                }
            }
        }
        if (!__any(live)) {
            if (disp.exhausted) break;
            continue;
        }
        // The bounce up to the light sample (pathBounce).
        bool cont = false, guided = false;
        Hit h;
        F3 qPos = mk3(0.f, 0.f, 0.f), qNrm = qPos, qWo = qPos;
        if (live) {
            if (recId != kNoRecord) {  // snapshot for the previous guided vertex's record
                rec[size_t(recId) * kRecF4 + 6] = make_float4(rad.r, rad.g, rad.b, 1.f);
                recId = kNoRecord;
            }
            bool done = false;
            if (closest(c, ray, h)) {
                Isect it = makeIsect(s, ray, h);
                const uint32_t mid = matIdOf(s, h.prim);
                const GMaterial& m = s.mats[mid];
                if (m.flags & 1u) {
                    rad = rad + emissionOf(m, mid, w) * beta;
                    done = true;
                } else {
                    BSDF bsdf = makeBSDF(s, m, mid, it, h, w);
                    if (MIS) {
                        rad = rad + evalMIS(c, it, bsdf, smp, w) * beta;
                    } else {
                        LightSample ls = sampleLight(c, it, bsdf, smp, w);
                        if (ls.pdf > 0.0f) rad = rad + ls.weight * beta / ls.pdf / ls.selPdf;
                    }
                    cont = true;
                    guided = bsdf.type != 2u && ((gmask >> bsdf.type) & 1u) &&
                             (guideDepth == 0u || depth < guideDepth);
                    qPos = it.pos;
                    qNrm = it.frame.N;
                    qWo = it.view;
                }
            } else {
                rad = rad + sp3(s.env) * beta;
                done = true;
            }
            if (done) finish();
        }
        const uint32_t gRank = guideWave<WD>(gv, guided, qPos, qNrm, qWo, tile, raw);
        // The continuation (kWfShadeCont's, with the guide's mixture).
        if (cont) {
            Isect it = makeIsect(s, ray, h);
            const uint32_t mid = matIdOf(s, h.prim);
            BSDF bsdf = makeBSDF(s, s.mats[mid], mid, it, h, w);
            F3 wi;
            float pdf = 0.0f, qmix = 0.f;
            Spec f;
            if (guided) {
                const RawMixture mix = rawMixture(raw + gRank * kGuideRaw);
                f = guidedSampleEval(bsdf, mix, alpha, smp, it.view, wi, pdf, qmix);
            } else {
                f = bsdfSampleEval(bsdf, smp, it.view, wi, pdf);
            }
            bool alive = true, record = false;
            if (pdf <= 0.0f) {
                alive = false;
            } else {
                beta = beta * (f / pdf);
                record = guided && rec != nullptr;
                // the lanes here all continue; those recording reserve slots
                const uint32_t slot = wavePush(recCounter, record);
                if (record) {
                    if (slot < recCap) {
                        float4* R = rec + size_t(slot) * kRecF4;
                        R[0] = make_float4(it.pos.x, it.pos.y, it.pos.z, qmix);
                        R[1] = make_float4(it.frame.N.x, it.frame.N.y, it.frame.N.z, 0.f);
                        R[2] = make_float4(it.view.x, it.view.y, it.view.z, 0.f);
                        R[3] = make_float4(wi.x, wi.y, wi.z, __uint_as_float(uint32_t(out)));
                        R[4] = make_float4(f.r, f.g, f.b, 0.f);
                        R[5] = make_float4(beta.r, beta.g, beta.b, 0.f);  // before roulette
                        R[6] = make_float4(0.f, 0.f, 0.f, 0.f);
                        recId = slot;
                    }
                }
                ray.o = offsetRayOrigin(it.pos, oriented(it));
                ray.d = wi;
                if (smp.next1D() < cfg.rrThreshold) alive = false;
                else if (!roulette(cfg, depth, smp, beta)) alive = false;
            }
            if (alive && ++depth >= cfg.maxDepth) alive = false;
            if (!alive) finish();
        }
    }
    unsigned long long r64 = c.rays;  // rays traced, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) r64 += __shfl_down(r64, off, 64);
    if (lane == 0 && r64) atomicAdd(counters, r64);
}

#if CRH_GMK_PART == 3
// Records -> training samples, as kWfRecordsToSamples with the final
// radiance from L.
__global__ __launch_bounds__(256) void kGuidedRecordsToSamples(const float4* __restrict__ rec,
                                                               const uint32_t* __restrict__ recCounter,
                                                               uint32_t recCap,
                                                               const float4* __restrict__ L,
                                                               GuideSample* __restrict__ out,
                                                               uint32_t cap,
                                                               uint32_t* __restrict__ trainCounter,
                                                               uint32_t radianceTarget) {
    const uint32_t n = min(*recCounter, recCap);
    for (uint32_t base = blockIdx.x * 256u; base < n; base += gridDim.x * 256u) {
        const uint32_t r = base + threadIdx.x;
        bool ok = false;
        float weight = 0.f;
        const float4* R = rec + size_t(r < n ? r : 0) * kRecF4;
        if (r < n) {
            const float4 snap = R[6];
            ok = snap.w == 1.f;
            if (ok) {
                const uint32_t p = __float_as_uint(R[3].w);
                const float4 e = L[p], f = R[4], bw = R[5];
                const float li[3] = {bw.x > 0.f ? (e.x - snap.x) / bw.x : 0.f,
                                     bw.y > 0.f ? (e.y - snap.y) / bw.y : 0.f,
                                     bw.z > 0.f ? (e.z - snap.z) / bw.z : 0.f};
                const float y = radianceTarget
                                    ? (li[0] + li[1] + li[2]) / 3.f
                                    : (f.x * li[0] + f.y * li[1] + f.z * li[2]) / 3.f;
                weight = y / R[0].w;
                ok = weight > 0.f && isfinite(weight);
            }
        }
        const uint32_t slot = wavePush(trainCounter, ok);
        if (ok && slot < cap) {
            GuideSample& g = out[slot];
            const float4 a = R[0], b = R[1], cc = R[2], d = R[3];
            g.pos[0] = a.x; g.pos[1] = a.y; g.pos[2] = a.z; g.pos[3] = 0.f;
            g.nrm[0] = b.x; g.nrm[1] = b.y; g.nrm[2] = b.z; g.nrm[3] = 0.f;
            g.wo[0] = cc.x; g.wo[1] = cc.y; g.wo[2] = cc.z; g.wo[3] = 0.f;
            g.wi[0] = d.x; g.wi[1] = d.y; g.wi[2] = d.z; g.wi[3] = weight;
        }
    }
}

__global__ void kGuidedReset(uint32_t* __restrict__ a, uint32_t* __restrict__ b) {
    if (threadIdx.x == 0) {
        *a = 0u;
        if (b) *b = 0u;
    }
}

#endif  // CRH_GMK_PART == 3

}  // namespace

#if CRH_GMK_PART == 3
size_t guidedScratchFloats(uint32_t grid) {
    return size_t(grid) * (kRenderBlock / 64u) * 64u * kGuideRaw;
}

size_t guidedLdsBytes(uint32_t stackMode, uint32_t stackCap, uint32_t width) {
    const size_t st = (stackLdsBytes(stackMode, stackCap, kRenderBlock) + 15) & ~size_t(15);
    return st + (kRenderBlock / 64u) * (width == 128 ? kTileBytesOf<128> : kTileBytesOf<64>);
}

hipError_t launchGuidedPass(const SceneView& s, const CameraView& cam, const IntegratorView& cfg,
                            uint32_t si, uint32_t active, float4* L, float2* P,
                            unsigned long long* counters, const Tuning& t, uint32_t* next,
                            const GuidedPassArgs& g, hipStream_t st, const GuidedViews* views) {
    const uint32_t tilesX = (cam.width + 15u) / 16u, tilesY = (cam.height + 15u) / 16u;
    const uint32_t tiles = views ? views->tiles : tilesX * tilesY;
    if (uint64_t(tiles) * active >= (uint64_t(1) << 23)) return hipErrorInvalidValue;
    if (!g.guide || !g.scratch || !next) return hipErrorInvalidValue;
    const uint32_t total = tiles * active * 256u;
    const uint32_t grid = std::min<uint32_t>(tiles * active, t.guidedGrid);
    const uint32_t tileOff =
        uint32_t((stackLdsBytes(t.guidedPriv, s.stackCap, kRenderBlock) + 15) & ~size_t(15));
    const size_t lds = guidedLdsBytes(t.guidedPriv, s.stackCap, g.guide->width);
    hipLaunchKernelGGL(kGuidedReset, dim3(1), dim3(64), 0, st, next, g.recCounter);
    const GuidedLaunch gl{s, cam, cfg, si, tilesX, tiles, total, next, L, P, counters, &g,
                          grid, tileOff, lds, t.guidedWaves, st,
                          views ? *views : GuidedViews{}, views != nullptr};
    hipError_t e = hipSuccess;
    withStack(t.guidedPriv, [&](auto sm) { e = launchGuidedSM<decltype(sm)::value>(gl); });
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t launchGuidedRecordsToSamples(const float4* rec, const uint32_t* recCounter,
                                        uint32_t recCap, const float4* L, GuideSample* out,
                                        uint32_t cap, uint32_t* trainCounter,
                                        uint32_t radianceTarget, hipStream_t st) {
    hipLaunchKernelGGL(kGuidedReset, dim3(1), dim3(64), 0, st, trainCounter, nullptr);
    hipLaunchKernelGGL(kGuidedRecordsToSamples, dim3(2048), dim3(256), 0, st, rec, recCounter,
                       recCap, L, out, cap, trainCounter, radianceTarget);
    return hipGetLastError();
}

#endif  // CRH_GMK_PART == 3

#if CRH_GMK_PART < 3
// kRenderGuided for one stack mode (this object's CRH_GMK_PART), every
// register budget and MIS on / off.
template <>
hipError_t launchGuidedSM<CRH_GMK_PART>(const GuidedLaunch& a) {
    constexpr int SM = CRH_GMK_PART;
    const GuidedPassArgs& g = *a.g;
    withWaves(a.waves, [&](auto wv) {
        constexpr int Wv = decltype(wv)::value;
        auto go = [&](auto wd) {
            constexpr int WD = decltype(wd)::value;
#define CRH_RG(MIS, MULTI)                                                                     \
    hipLaunchKernelGGL((kRenderGuided<MIS, SM, Wv, MULTI, WD>), dim3(a.grid),                  \
                       dim3(kRenderBlock), a.lds, a.st, a.s, a.cam, a.cfg, a.si, a.tilesX,       \
                       a.tiles, a.total, a.next, a.L, a.P, a.counters, *g.guide, g.alpha,        \
                       g.mask, g.maxDepth, g.scratch, a.tileOff, g.rec, g.recCounter, g.recCap, \
                       a.views)
            if (a.cfg.useMIS && a.multi) CRH_RG(true, true);
            else if (a.multi) CRH_RG(false, true);
            else if (a.cfg.useMIS) CRH_RG(true, false);
            else CRH_RG(false, false);
#undef CRH_RG
        };
        if (g.guide->width == 128) go(std::integral_constant<int, 128>{});
        else go(std::integral_constant<int, 64>{});
    });
    return hipGetLastError();
}
#endif

}  // namespace crh
