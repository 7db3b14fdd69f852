// Layout of the per-scene spectral transport table (host and device), in
// floats: the film's CIE weights (x, y, z, pad per bin), the environment
// spectrum, the averages and the rows of the six coloured basis spectra of
// the RGB uplift (cyan, magenta, yellow, red, green, blue; for textured
// materials), then one block per material: the averages of its diffuse and
// specular spectra over all bins, then its diffuse, specular and emission
// spectra.  Bins are 5 nm wide from 360 nm to 830 nm.
#pragma once

namespace crh {

constexpr int kSpecBins = 95;
constexpr int kSpecCmf = 0;
constexpr int kSpecEnv = 4 * kSpecBins;  // 380
constexpr int kSpecBasisAvg = kSpecEnv + 96;         // 476: 6 averages + 2 pad
constexpr int kSpecBasis = kSpecBasisAvg + 8;         // 484: 6 x 95 bins
constexpr int kSpecMat0 = kSpecBasis + 6 * kSpecBins + 2;  // 1056
constexpr int kSpecMatStride = 296;
constexpr int kSpecAvgKd = 0, kSpecAvgKs = 1, kSpecKd = 8, kSpecKs = 8 + kSpecBins,
              kSpecEm = 8 + 2 * kSpecBins;

}  // namespace crh
