"""Generates cie_table.h: the CIE 1931 2-degree colour matching functions at the
95 spectral bins, with the step and Y normalisation the XYZ conversion uses.

The values are the reference's kXYZTable (Core/SpectrumTabData.h:7-55), which
its Tools/spectrum.py made from the CIE's 1 nm table (Tools/CIE_xyz_1931_2deg.csv,
rows 360..830 nm): it skips the first row (the file has no header, so the
360 nm row goes), samples the remaining 361..830 nm range at 95 evenly spaced
wavelengths (step 469/94 nm) by linear interpolation, and prints 6 decimals.
This script repeats those steps so that the table is the reference's number
for number; kXYZStep = 470/95 (Core/SpectrumTabData.h:57) and
kCIENormalizationY = 1/106.8569131841719 (Core/Spectrum.h:286) are carried
as written there.

Run: python craystal_amd/csrc/make_cie_table.py [path/to/CIE_xyz_1931_2deg.csv]
(default: the reference checkout's Tools/CIE_xyz_1931_2deg.csv); rewrites
cie_table.h next to this file.
"""
from __future__ import annotations

import sys
from pathlib import Path

BINS = 95
DEFAULT_CSV = Path("/root/reference/Tools/CIE_xyz_1931_2deg.csv")


def resample(rows: list[list[float]]) -> list[list[float]]:
    """(3, 95) table from the CSV rows (lambda, x, y, z), as the reference's
    generator does it (its first row is dropped as if it were a header)."""
    rows = rows[1:]
    lam = [r[0] for r in rows]
    step = (lam[-1] - lam[0]) / (BINS - 1)
    out = [[], [], []]
    for k in range(BINS):
        w = lam[0] + k * step
        for i in range(len(lam) - 1):
            if lam[i] <= w <= lam[i + 1]:
                t = (w - lam[i]) / (lam[i + 1] - lam[i])
                for c in range(3):
                    out[c].append(rows[i][c + 1] * (1 - t) + rows[i + 1][c + 1] * t)
                break
    assert all(len(o) == BINS for o in out)
    return out


def main() -> None:
    csv = Path(sys.argv[1]) if len(sys.argv) > 1 else DEFAULT_CSV
    rows = [list(map(float, ln.strip().split(","))) for ln in csv.read_text().splitlines()
            if ln.strip()]
    tab = resample(rows)
    names = ("x-bar", "y-bar", "z-bar")
    out = ["// Generated by make_cie_table.py from the CIE 1931 2-degree observer (1 nm",
           "// table); do not edit.  Bin k holds the colour matching functions at",
           "// 361 + k * 469/94 nm, 6 decimals (the reference's kXYZTable,",
           "// Core/SpectrumTabData.h:7-55).",
           "#pragma once", "", "namespace crh {", "",
           "constexpr int kCieTableBins = 95;",
           "// Step and Y normalisation of the XYZ sum (Core/SpectrumTabData.h:57,",
           "// Core/Spectrum.h:286): XYZ = sum_k cmf_k * s_k * kCieXYZStep * kCieNormY.",
           "constexpr float kCieXYZStep = (830.0f - 360.0f) / float(kCieTableBins);",
           "constexpr float kCieNormY = float(1.0 / 106.8569131841719);", "",
           "constexpr float kCieXYZ[3][95] = {"]
    for name, row in zip(names, tab):
        vals = [f"{v:.6f}f" for v in row]
        lines = [", ".join(vals[i:i + 6]) for i in range(0, BINS, 6)]
        out.append(f"    {{// {name}")
        out += [f"     {ln}," for ln in lines[:-1]] + [f"     {lines[-1]}}},"]
    out += ["};", "", "}  // namespace crh", ""]
    Path(__file__).with_name("cie_table.h").write_text("\n".join(out))


if __name__ == "__main__":
    main()
