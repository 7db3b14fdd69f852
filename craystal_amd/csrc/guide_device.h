// Device-side use of the guide's vMF mixture in the integrator: density,
// sampling (W. Jakob, "Numerically stable sampling of the von Mises-Fisher
// distribution on S^2", 2012) and the actual density of the reference's
// BSDF sampling procedures, which the guided estimator needs to keep the
// plain estimator's expectation (see kWfShadeCont).
#pragma once
#include "shading.h"

namespace crh {

struct Mixture {
    float w[8];
    F3 mu[8];
    float k[8];
};

CRH_D Mixture loadMixture(const float* p) {
    Mixture m;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        m.w[j] = p[j];
        m.mu[j] = mk3(p[8 + 3 * j], p[9 + 3 * j], p[10 + 3 * j]);
        m.k[j] = p[32 + j];
    }
    return m;
}

CRH_D float vmfDensity(float k, float c) {
    return k / (2.f * kPi * (-expm1f(-2.f * k))) * expf(k * (c - 1.f));
}

CRH_D float mixturePdf(const Mixture& m, F3 d) {
    float q = 0.f;
#pragma unroll
    for (int j = 0; j < 8; ++j) q += m.w[j] * vmfDensity(m.k[j], dot(m.mu[j], d));
    return q;
}

// Lobe chosen by u0 against the weights' CDF; direction from (u1, u2).
CRH_D F3 mixtureSample(const Mixture& m, float u0, float u1, float u2) {
    int j = 7;
    float c = 0.f;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        c += m.w[i];
        if (u0 < c) {
            j = i;
            break;
        }
    }
    float k = m.k[j];
    F3 mu = m.mu[j];
    float w = 1.f + logf(u1 + (1.f - u1) * expf(-2.f * k)) / k;
    w = fminf(fmaxf(w, -1.f), 1.f);
    float s = sqrtf(fmaxf(0.f, 1.f - w * w));
    F3 a = fabsf(mu.x) > 0.9f ? mk3(0, 1, 0) : mk3(1, 0, 0);
    F3 t = normalize(cross(a, mu));
    F3 b = cross(mu, t);
    float phi = 2.f * kPi * u2;
    return normalize(mu * w + (t * cosf(phi) + b * sinf(phi)) * s);
}

// Density of the directions the reference's sampling procedure produces
// (Material/*BRDF.cu sampleImpl), which differs from the pdf it reports
// for the Principled lobe mixture.  Lambertian: cosine.  Principled: lobe
// chosen with probability pS, GGX half-vector density D(h) cos(theta_h)
// mapped by the reflection Jacobian 1 / (4 |wo.h|), else cosine.
CRH_D float actualPdfLocal(const BSDF& b, F3 wo, F3 wi) {
    if (b.type == 0u) return cosHemiPdf(wi);
    if (b.type == 1u) {
        if (wo.z <= 0.0f) return 0.f;
        float spec = 0.f;
        // The half vector the sampler drew is the one of +-(wo + wi) with
        // positive z (the other sign maps to the same wi, below the horizon).
        F3 h = normalize(wo + wi);
        if (h.z < 0.f) h = h * -1.f;
        float cd = dot(wo, h);
        if (h.z > 0.f && fabsf(cd) > 0.f) spec = ggxD(b.rough, h.z) * h.z / (4.f * fabsf(cd));
        return b.pS * spec + (1.f - b.pS) * cosHemiPdf(wi);
    }
    return 0.f;  // Phong: not guided
}

}  // namespace crh
