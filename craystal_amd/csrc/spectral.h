// Host side of spectral transport: the CIE 1931 colour matching functions
// and the per-scene spectral table the megakernel reads (layout: kSpec* in
// shading.h).
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace crh {

constexpr int kCieBins = 95;            // 360, 365, ..., 830 nm
constexpr double kCieLambda0 = 360.0;
constexpr double kCieStep = 5.0;

// x-bar, y-bar, z-bar at lambda (nm): the multi-lobe piecewise-Gaussian fit of
// C. Wyman, P.-P. Sloan, P. Shirley, "Simple Analytic Approximations to the
// CIE XYZ Color Matching Functions", JCGT 2(2), 2013.
void cieFit(double lambda, double xyz[3]);

// Normalisation so that a constant spectrum 1 has Y = 1:
// 1 / (sum_i y-bar(lambda_i) * step).
double cieYNorm();

// XYZ of a 95-bin spectrum: sum_i (x, y, z)-bar(lambda_i) * s_i * step * norm.
void spectrumToXYZ(const float* s, float xyz[3]);

// RGB -> 95-bin spectrum (Smits's decomposition into white + one secondary +
// one primary basis spectrum, whose weights are the sorted RGB differences).
// Linear in rgb within each ordering of the channels and homogeneous, so it
// serves reflectances (0..1 in, 0..1 out) and emission alike.  Negative
// channels are taken as 0.
void rgbToSpectrum(const float rgb[3], float out[kCieBins]);

// The table for a scene: film CIE weights, environment spectrum, then per
// material the averages of its diffuse / specular spectra and the three
// spectra.  spectra: nmaterial x 3 x 95 floats (diffuse, specular, emission).
std::vector<float> buildSpectralTable(const float* spectra, uint32_t nmaterial,
                                      const float* envSpectrum);
std::vector<float> upliftedSpectra(const float* rgb9, uint32_t nmaterial, size_t stride);

}  // namespace crh
