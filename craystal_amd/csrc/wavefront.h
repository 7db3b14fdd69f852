// Wavefront integrator state (structure of arrays in HBM, one slot per path
// of a pass) and its host entry point.
#pragma once
#include <hip/hip_runtime.h>

#include "layout.h"

namespace crh {

enum : uint32_t {
    kCExt0 = 0,  // extension queue, ping
    kCExt1 = 1,  // extension queue, pong
    kCMis0 = 2,  // MIS-ray queue, ping
    kCMis1 = 3,  // MIS-ray queue, pong
    kCShadow = 4,
    kCCont = 5,  // paths to shade after the shadow query
    kCRec = 6,   // training records written this pass
    kCTrain = 7, // training samples produced from the records
    kCPtExt = 8,     // ray hand-out counter of the persistent extend kernel
    kCPtShadow = 9,  // ray hand-out counter of the persistent shadow kernel
    kWfCounters = 10,
};

struct WfState {
    uint32_t width, height;
    uint32_t npix;    // width * height
    uint32_t slots;   // samples per pixel the state holds (paths = slots * npix)
    uint32_t active;  // samples per pixel of the current pass group
    float4* rayO;  // extension ray: origin, tMin
    float4* rayD;  // direction, tMax
    uint4* hit;    // closest hit of the extension ray
    float4* beta;
    float4* rad;   // radiance; path p = slot * npix + pixel
    uint32_t* rng; // Sampler state (the increment is recomputed from the pixel)
    uint32_t* flags;
    float2* P;     // jittered film position
    float4* shO;   // shadow ray
    float4* shD;
    float4* lgt;   // light direction, d^2
    float4* lgn;   // interpolated light normal, emitter area
    uint32_t* lgId;
    float4* misO;  // MIS BSDF ray
    float4* misD;
    uint4* misHit;
    float4* misVal;  // running MIS value, selection pdf
    float4* misBw;   // BSDF weight, BSDF pdf
    float4* misPos;  // shading position, emitter id bits
    float4* misBeta; // path throughput at the MIS vertex
    uint32_t* qExt[2];
    uint32_t* qMis[2];
    uint32_t* qShadow;
    uint32_t* qCont;
    uint32_t* counters;
    unsigned long long* work;  // rays, node visits, triangle tests
    // Guided mode only (null otherwise).
    float4* gPos;     // shading point of the vertex being continued
    float4* gNrm;     // shading normal
    float4* gWo;      // outgoing (view) direction
    float* gMix;      // 40 floats per path: mixture weights, means, kappas
    float4* rec;      // training records, kRecF4 float4 each
    uint32_t* recId;  // per path: record of its last guided vertex
    uint32_t recCap;
};

constexpr uint32_t kRecF4 = 7;  // float4s per training record
constexpr uint32_t kNoRecord = 0xFFFFFFFFu;

struct WfLaunch {
    uint32_t traceGrid = 2048;   // blocks for the traversal kernels
    uint32_t traceBlock = 256;   // threads per traversal block (64..256)
    uint32_t shadeGrid = 2048;   // blocks of 256 for the shading kernels
    uint32_t privStack = 0;      // stack mode: 0 LDS, 1 private, 2 short LDS + private
    uint32_t persistent = 0;     // 1: traversal lanes refill with new rays
    uint32_t refill = 16;
    uint32_t finishDepth = 0;    // > 0: bounces from this depth on run per path in one launch
    uint32_t finishWaves = 4;    // register budget of that launch (1 or 4 waves per SIMD)
    uint32_t finishGrid = 2048;  // its persistent blocks of 256
};

struct GuideView;
struct GuideSample;

// spectral: hero-wavelength transport of the scene's spectra (not with a
// guide or counted work); the pass's radiance buffer then holds XYZ.
// guide != null: guided continuation sampling (one-sample MIS with the
// guide's mixture, fraction alpha for BSDF sampling).  samples != null:
// also convert this pass's vertex records into training samples.
// Renders sample indices si .. si + active - 1 of every pixel at once (one
// path per pixel and sample, active <= w.slots); the caller gathers the
// slots into the film in sample order.
hipError_t wavefrontPass(const SceneView& s, const CameraView& cam,
                         const IntegratorView& cfg, uint32_t si, uint32_t active, WfState& w,
                         const WfLaunch& L, bool countWork, hipStream_t st,
                         const GuideView* guide = nullptr, float alpha = 0.5f,
                         GuideSample* samples = nullptr, uint32_t sampleCap = 0,
                         uint32_t guideMask = 3u, uint32_t guideDepth = 0u,
                         uint32_t guideTarget = 0u, bool spectral = false);

// One pass of a batch: sample indices si .. si + active - 1 of every pixel
// of the first nActive scenes of the table (pass pixel q is pixel
// q - pixPrefix[j] of scene j, pixPrefix[nActive] = w.npix), their films'
// samples in w.rad / w.P at slot * w.npix + q.  Scene views and cameras come
// from the table; stackCap: the deepest of their trees.  No guide, no
// counted work.
hipError_t wavefrontBatchPass(const BatchScene* tab, const uint32_t* pixPrefix, uint32_t nActive,
                              uint32_t stackCap, const IntegratorView& cfg, uint32_t si,
                              uint32_t active, WfState& w, const WfLaunch& L, hipStream_t st,
                              bool spectral);

// The finishing launch (kWfFinish) of one pass; launchWfFinishSM<SM> is
// defined once per stack mode (wf_finish.hip built with CRH_WF_SM = SM).
struct WfFinishArgs {
    SceneView s;  // one scene, or (batch) the table below
    WfState w;
    IntegratorView cfg;
    uint32_t cur, d0, grid;
    size_t lds;
    uint32_t waves;
    bool spectral;
    hipStream_t st;
    bool batch;
    const BatchScene* tab;
    const uint32_t* pixPrefix;
    uint32_t nActive;
};
template <int SM> hipError_t launchWfFinishSM(const WfFinishArgs& a);
template <> hipError_t launchWfFinishSM<0>(const WfFinishArgs&);
template <> hipError_t launchWfFinishSM<1>(const WfFinishArgs&);
template <> hipError_t launchWfFinishSM<2>(const WfFinishArgs&);

}  // namespace crh
