// Host-side view of the guiding network's device buffers and its launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "guiding.h"

namespace crh {

struct GuideView {
    float* theta;  // fp32 master parameters (gParamsOf(width))
    float* m;      // Adam first moment
    float* v;      // Adam second moment
    float* grad;   // gradient accumulator (cleared by the Adam kernel)
    uint16_t* wq;  // bf16 bits: W_l in MFMA B-operand order (gQElemsOf(width))
    uint16_t* wqT; // bf16 bits: W_l^T, rows padded to a multiple of 64, same order
    float bmin[3];
    float bext[3];
    uint32_t width;  // hidden width: 64 or 128
};

hipError_t guideSetLdsLimits();
hipError_t launchGuideInfer(const GuideView& gv, const uint32_t* queue, const uint32_t* count,
                            uint32_t countDirect, const float4* pos, const float4* nrm,
                            const float4* wo, float* params, float* rawOut, uint32_t grid,
                            hipStream_t st);
hipError_t launchGuideQuantize(const GuideView& gv, hipStream_t st);
hipError_t launchGuideTrainStep(const GuideView& gv, const GuideSample* samples, uint32_t n,
                                uint16_t* Hsave, uint16_t* Dsave, float* lossSum, float lr,
                                uint32_t step, bool applyAdam, hipStream_t st);

}  // namespace crh
