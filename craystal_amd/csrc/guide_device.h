// Device-side use of the guide's vMF mixture in the integrator: density,
// sampling (W. Jakob, "Numerically stable sampling of the von Mises-Fisher
// distribution on S^2", 2012) and the actual density of the reference's
// BSDF sampling procedures, which the guided estimator needs to keep the
// plain estimator's expectation (see kWfShadeCont).
#pragma once
#include "guide_mlp.h"
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

CRH_D float vmfDensity(float k, float c) { return vmfPdf(k, c); }

CRH_D float mixturePdf(const Mixture& m, F3 d) {
    float q = 0.f;
#pragma unroll
    for (int j = 0; j < 8; ++j) q += m.w[j] * vmfDensity(m.k[j], dot(m.mu[j], d));
    return q;
}

// A direction from the vMF lobe (k, mu) by W. Jakob's stable inversion
// (2012) from (u1, u2).
CRH_D F3 vmfSampleDir(float k, F3 mu, float u1, float u2) {
    float w = 1.f + logf(u1 + (1.f - u1) * expf(-2.f * k)) / k;
    w = fminf(fmaxf(w, -1.f), 1.f);
    float s = sqrtf(fmaxf(0.f, 1.f - w * w));
    F3 a = fabsf(mu.x) > 0.9f ? mk3(0, 1, 0) : mk3(1, 0, 0);
    F3 t = normalize(cross(a, mu));
    F3 b = cross(mu, t);
    float phi = 2.f * kPi * u2;
    return normalize(mu * w + (t * cosf(phi) + b * sinf(phi)) * s);
}

// Lobe chosen by u0 against the weights' CDF (the last one if rounding
// leaves u0 above it), selected with static indices (no early exit) so the
// mixture stays in registers; direction from (u1, u2).
CRH_D F3 mixtureSample(const Mixture& m, float u0, float u1, float u2) {
    float k = m.k[7];
    F3 mu = m.mu[7];
    float c = 0.f;
    bool found = false;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        c += m.w[i];
        const bool take = !found && u0 < c;
        k = take ? m.k[i] : k;
        mu = take ? m.mu[i] : mu;
        found = found || take;
    }
    return vmfSampleDir(k, mu, u1, u2);
}

// The same mixture read lobe by lobe from the network's raw outputs o (the
// head's helpers, so every value equals computeHead's), for kernels that
// cannot afford to hold 40 parameters in registers.
struct RawMixture {
    const float* o;
    float amax, sum;
};
CRH_D RawMixture rawMixture(const float* o) {
    const float amax = headLogitMax(o);
    return RawMixture{o, amax, headExpSum(o, amax)};
}
CRH_D F3 rawMean(const RawMixture& m, int j) {
    const float il = gRcp(headMeanLength(m.o, j));
    return mk3(m.o[5 * j + 1] * il, m.o[5 * j + 2] * il, m.o[5 * j + 3] * il);
}
CRH_D float mixturePdf(const RawMixture& m, F3 d) {
    float q = 0.f;
#pragma unroll
    for (int j = 0; j < 8; ++j)
        q += headWeight(m.o, j, m.amax, m.sum) * vmfDensity(headKappa(m.o, j), dot(rawMean(m, j), d));
    return q;
}
CRH_D F3 mixtureSample(const RawMixture& m, float u0, float u1, float u2) {
    int sel = 7;
    float c = 0.f;
    bool found = false;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        c += headWeight(m.o, i, m.amax, m.sum);
        const bool take = !found && u0 < c;
        sel = take ? i : sel;
        found = found || take;
    }
    return vmfSampleDir(headKappa(m.o, sel), rawMean(m, sel), u1, u2);
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

// One-sample MIS of BSDF sampling (fraction alpha) and the guide's mixture
// at a vertex: draws the technique, then the BSDF's 2D sample or the lobe
// and direction of the mixture.  Returns f * |cos| at wi; pdf is set so that
// the weight f / pdf = f * (p_act / p_rep) / q_mix, whose expectation is the
// plain estimator's f / p_rep with wi ~ p_act, so guided and plain images
// converge to the same (reference) result; p_rep is the pdf the reference
// reports, p_act the density its sampler actually has, and
// q_mix = alpha p_act + (1 - alpha) q_guide (returned in qmix).
template <class Mix>
CRH_D Spec guidedSampleEval(const BSDF& bsdf, const Mix& mix, float alpha, Sampler& smp,
                            F3 view, F3& wi, float& pdf, float& qmix) {
    F3 woL = toLocal(bsdf.frame, view);
    F3 wiL;
    if (smp.next1D() < alpha) {
        F2 u = smp.next2D();
        float dummy;
        wiL = sampleLocal(bsdf, woL, u, dummy);
        wi = toWorld(bsdf.frame, wiL);
    } else {
        float u0 = smp.next1D();
        F2 u = smp.next2D();
        wi = mixtureSample(mix, u0, u.x, u.y);
        wiL = toLocal(bsdf.frame, wi);
    }
    Spec f = evalLocal(bsdf, woL, wiL);
    float prep = pdfLocal(bsdf, woL, wiL);
    float pact = actualPdfLocal(bsdf, woL, wiL);
    qmix = alpha * pact + (1.f - alpha) * mixturePdf(mix, wi);
    pdf = (prep > 0.f && pact > 0.f && qmix > 0.f) ? qmix * prep / pact : 0.f;
    return f;
}

}  // namespace crh
