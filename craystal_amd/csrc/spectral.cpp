#include "spectral.h"

#include "cie_table.h"
#include "rgb_basis.h"
#include "spec_layout.h"

#include <algorithm>
#include <cmath>

namespace crh {

namespace {
constexpr int kCmf = kSpecCmf, kEnv = kSpecEnv, kMat0 = kSpecMat0, kMatStride = kSpecMatStride;
constexpr int kAvgKd = kSpecAvgKd, kAvgKs = kSpecAvgKs, kKd = kSpecKd, kKs = kSpecKs,
              kEm = kSpecEm;
static_assert(kSpecBins == kCieBins && kCieTableBins == kCieBins, "bin count");

float average(const float* v) {
    float s = 0.0f;
    for (int i = 0; i < kCieBins; ++i) s += v[i];
    return s / float(kCieBins);
}
}  // namespace

void cieTable(int k, float xyz[3]) {
    for (int j = 0; j < 3; ++j) xyz[j] = kCieXYZ[j][k];
}

void spectrumToXYZ(const float* s, float xyz[3]) {
    float acc[3] = {0.0f, 0.0f, 0.0f};
    for (int i = 0; i < kCieBins; ++i)
        for (int j = 0; j < 3; ++j) acc[j] += kCieXYZ[j][i] * s[i];
    const float k = kCieNormY * kCieXYZStep;
    for (int j = 0; j < 3; ++j) xyz[j] = acc[j] * k;
}

void rgbToSpectrum(const float rgb[3], float out[kCieBins]) {
    enum { kC, kM, kY, kR, kG, kB };
    const float r = std::max(rgb[0], 0.0f), g = std::max(rgb[1], 0.0f),
                b = std::max(rgb[2], 0.0f);
    float white, w1, w2;
    int b1, b2;
    if (r <= g && r <= b) {  // red smallest: white + cyan + (blue or green)
        white = r;
        b1 = kC;
        if (g <= b) { w1 = g - r; b2 = kB; w2 = b - g; }
        else        { w1 = b - r; b2 = kG; w2 = g - b; }
    } else if (g <= r && g <= b) {  // green smallest: white + magenta + ...
        white = g;
        b1 = kM;
        if (r <= b) { w1 = r - g; b2 = kB; w2 = b - r; }
        else        { w1 = b - g; b2 = kR; w2 = r - b; }
    } else {  // blue smallest: white + yellow + (green or red)
        white = b;
        b1 = kY;
        if (r <= g) { w1 = r - b; b2 = kG; w2 = g - r; }
        else        { w1 = g - b; b2 = kR; w2 = r - g; }
    }
    for (int i = 0; i < kCieBins; ++i)
        out[i] = white + w1 * kRgbBasis[b1][i] + w2 * kRgbBasis[b2][i];
}

std::vector<float> upliftedSpectra(const float* rgb9, uint32_t nmaterial, size_t stride) {
    std::vector<float> s(size_t(nmaterial) * 3 * kCieBins);
    for (uint32_t m = 0; m < nmaterial; ++m)
        for (int k = 0; k < 3; ++k)
            rgbToSpectrum(rgb9 + m * stride + 3 * k, s.data() + (size_t(m) * 3 + k) * kCieBins);
    return s;
}

std::vector<float> buildSpectralTable(const float* spectra, uint32_t nmaterial,
                                      const float* envSpectrum) {
    std::vector<float> t(size_t(kMat0) + size_t(nmaterial) * kMatStride, 0.0f);
    // A path carries three of the 95 bins, each uniform: the film weight of a
    // lane is (x, y, z)-bar * step * norm * 95 / 3.
    const double k = double(kCieNormY * kCieXYZStep) * double(kCieBins) / 3.0;
    for (int i = 0; i < kCieBins; ++i)
        for (int j = 0; j < 3; ++j) t[kCmf + 4 * i + j] = float(double(kCieXYZ[j][i]) * k);
    if (envSpectrum)
        for (int i = 0; i < kCieBins; ++i) t[kEnv + i] = envSpectrum[i];
    for (int b = 0; b < 6; ++b) {
        for (int i = 0; i < kCieBins; ++i) t[kSpecBasis + b * kCieBins + i] = kRgbBasis[b][i];
        t[kSpecBasisAvg + b] = average(kRgbBasis[b]);
    }
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
