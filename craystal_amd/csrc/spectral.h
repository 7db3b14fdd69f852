// Host side of spectral transport: the CIE 1931 colour matching functions
// and the per-scene spectral table the megakernel reads (layout: kSpec* in
// spec_layout.h).
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace crh {

constexpr int kCieBins = 95;  // spectral bins; a spectrum's bin k pairs with table entry k
constexpr double kCieLambda0 = 360.0;  // nominal bin wavelengths 360, 365, ..., 830 nm
constexpr double kCieStep = 5.0;       // (used to sample smooth curves into bins)

// x-bar, y-bar, z-bar of bin k: the CIE 1931 2-degree table of the reference
// (kXYZTable, Core/SpectrumTabData.h:7-55; generated into cie_table.h).
void cieTable(int k, float xyz[3]);

// XYZ of a 95-bin spectrum as BroadSpectrum::toXYZ (Core/Spectrum.cpp:37-45)
// computes it: a float sum over bins of (x, y, z)-bar * s_k, then times
// kCIENormalizationY * kXYZStep (Core/Spectrum.h:286, SpectrumTabData.h:57).
// The reference reads x-bar for Z (Core/Spectrum.cpp:40); here Z uses z-bar.
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
