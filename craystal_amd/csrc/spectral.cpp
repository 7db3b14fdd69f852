#include "spectral.h"

#include "spec_layout.h"

#include <cmath>

namespace crh {

namespace {
// Piecewise Gaussian: sigma1 left of the mean, sigma2 right of it.
double lobe(double x, double mu, double s1, double s2) {
    double t = (x - mu) / (x < mu ? s1 : s2);
    return std::exp(-0.5 * t * t);
}

constexpr int kCmf = kSpecCmf, kEnv = kSpecEnv, kMat0 = kSpecMat0, kMatStride = kSpecMatStride;
constexpr int kAvgKd = kSpecAvgKd, kAvgKs = kSpecAvgKs, kKd = kSpecKd, kKs = kSpecKs,
              kEm = kSpecEm;
static_assert(kSpecBins == kCieBins, "bin count");

float average(const float* v) {
    float s = 0.0f;
    for (int i = 0; i < kCieBins; ++i) s += v[i];
    return s / float(kCieBins);
}
}  // namespace

void cieFit(double l, double xyz[3]) {
    xyz[0] = 1.056 * lobe(l, 599.8, 37.9, 31.0) + 0.362 * lobe(l, 442.0, 16.0, 26.7) -
             0.065 * lobe(l, 501.1, 20.4, 26.2);
    xyz[1] = 0.821 * lobe(l, 568.8, 46.9, 40.5) + 0.286 * lobe(l, 530.9, 16.3, 31.1);
    xyz[2] = 1.217 * lobe(l, 437.0, 11.8, 36.0) + 0.681 * lobe(l, 459.0, 26.0, 13.8);
}

double cieYNorm() {
    double sum = 0.0, c[3];
    for (int i = 0; i < kCieBins; ++i) {
        cieFit(kCieLambda0 + kCieStep * i, c);
        sum += c[1] * kCieStep;
    }
    return 1.0 / sum;
}

void spectrumToXYZ(const float* s, float xyz[3]) {
    const double k = cieYNorm();
    double acc[3] = {0.0, 0.0, 0.0}, c[3];
    for (int i = 0; i < kCieBins; ++i) {
        cieFit(kCieLambda0 + kCieStep * i, c);
        for (int j = 0; j < 3; ++j) acc[j] += c[j] * double(s[i]) * kCieStep * k;
    }
    for (int j = 0; j < 3; ++j) xyz[j] = float(acc[j]);
}

std::vector<float> buildSpectralTable(const float* spectra, uint32_t nmaterial,
                                      const float* envSpectrum) {
    std::vector<float> t(size_t(kMat0) + size_t(nmaterial) * kMatStride, 0.0f);
    // A path carries three of the 95 bins, each uniform: the film weight of a
    // lane is (x, y, z)-bar * step * norm * 95 / 3.
    const double k = cieYNorm() * kCieStep * double(kCieBins) / 3.0;
    double c[3];
    for (int i = 0; i < kCieBins; ++i) {
        cieFit(kCieLambda0 + kCieStep * i, c);
        for (int j = 0; j < 3; ++j) t[kCmf + 4 * i + j] = float(c[j] * k);
    }
    if (envSpectrum)
        for (int i = 0; i < kCieBins; ++i) t[kEnv + i] = envSpectrum[i];
    for (uint32_t m = 0; m < nmaterial; ++m) {
        float* dst = t.data() + kMat0 + size_t(m) * kMatStride;
        const float* src = spectra + size_t(m) * 3 * kCieBins;
        for (int i = 0; i < kCieBins; ++i) {
            dst[kKd + i] = src[i];
            dst[kKs + i] = src[kCieBins + i];
            dst[kEm + i] = src[2 * kCieBins + i];
        }
        dst[kAvgKd] = average(src);
        dst[kAvgKs] = average(src + kCieBins);
    }
    return t;
}

}  // namespace crh
