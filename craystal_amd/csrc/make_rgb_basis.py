"""Generates rgb_basis.h: the six coloured basis spectra of the RGB -> 95-bin
spectrum uplift (B. Smits, "An RGB-to-Spectrum Conversion for Reflectances",
Journal of Graphics Tools 4(4), 1999).  The reference declares this
conversion (BroadSpectrum::fromRGB, Core/Spectrum.cpp:49) but returns zeros;
its Smits tables (Core/SpectrumTabData.h:67-257) are not used here.

Smits's construction, recomputed for this library's colour pipeline: each
basis spectrum s (95 bins, 360..830 nm in 5 nm steps) is the smoothest
curve (least sum of squared neighbour differences) with 0 <= s <= 1 whose
colour is the target exactly:

    RGB_E(s) = diag(1 / w) * M * XYZ(s)

M is the XYZ -> linear sRGB matrix (the inverse of the sRGB -> XYZ matrix
the film uses), XYZ(s) the sum of the CIE table (cie_table.h, the
reference's kXYZTable) times s, and w = M * XYZ(1) so that the flat spectrum 1 has
RGB_E = (1, 1, 1) (white balanced to the equal-energy spectrum).  White
itself is the flat spectrum 1; targets: cyan (0,1,1), magenta (1,0,1),
yellow (1,1,0), red (1,0,0), green (0,1,0), blue (0,0,1).

Run: python craystal_amd/csrc/make_rgb_basis.py  (rewrites rgb_basis.h)
"""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np
from scipy.optimize import minimize

BINS = 95
SRGB_TO_XYZ = np.array([[0.412453, 0.357580, 0.180423],
                        [0.212671, 0.715160, 0.072169],
                        [0.019334, 0.119193, 0.950227]])
TARGETS = [("cyan", (0, 1, 1)), ("magenta", (1, 0, 1)), ("yellow", (1, 1, 0)),
           ("red", (1, 0, 0)), ("green", (0, 1, 0)), ("blue", (0, 0, 1))]


def cie_table() -> np.ndarray:
    """(3, 95): the colour matching functions of the bins, read from cie_table.h."""
    text = Path(__file__).with_name("cie_table.h").read_text()
    body = text[text.index("kCieXYZ[3][95] = {"):]
    vals = [float(v) for v in re.findall(r"(-?[0-9]+\.[0-9]+)f", body)]
    assert len(vals) == 3 * BINS
    return np.array(vals).reshape(3, BINS)


def rgb_e_matrix() -> np.ndarray:
    """(3, 95): spectrum -> white-balanced linear sRGB."""
    to_xyz = cie_table()                                # the scale divides out below
    m = np.linalg.inv(SRGB_TO_XYZ) @ to_xyz
    w = m.sum(axis=1)                                   # RGB of the flat spectrum 1
    return m / w[:, None]


def solve(target, A) -> np.ndarray:
    D = np.diff(np.eye(BINS), axis=0)
    Q = D.T @ D
    t = np.asarray(target, float)
    res = minimize(lambda s: s @ Q @ s, np.full(BINS, 0.5), jac=lambda s: 2 * Q @ s,
                   constraints=[{"type": "eq", "fun": lambda s: A @ s - t, "jac": lambda s: A}],
                   bounds=[(0.0, 1.0)] * BINS, method="SLSQP",
                   options={"maxiter": 2000, "ftol": 1e-15})
    s = np.clip(res.x, 0.0, 1.0)
    err = np.abs(A @ s - t).max()
    if not res.success or err > 1e-6:
        raise RuntimeError(f"basis {target}: {res.message}, residual {err}")
    return s


def main() -> None:
    A = rgb_e_matrix()
    rows = [(name, solve(t, A)) for name, t in TARGETS]
    out = ["// Generated by make_rgb_basis.py (Smits's RGB -> spectrum basis, recomputed for",
           "// this library's CIE table and sRGB matrix); do not edit.",
           "#pragma once", "", "namespace crh {", "",
           "// Basis spectra, 95 bins from 360 nm in 5 nm steps; white is the flat 1.",
           "// Rows: cyan, magenta, yellow, red, green, blue.",
           "constexpr float kRgbBasis[6][95] = {"]
    for name, s in rows:
        vals = [f"{v:.7f}f" for v in s]
        lines = [", ".join(vals[i:i + 6]) for i in range(0, BINS, 6)]
        out.append(f"    {{// {name}")
        out += [f"     {ln}," for ln in lines[:-1]] + [f"     {lines[-1]}}},"]
    out += ["};", "", "}  // namespace crh", ""]
    Path(__file__).with_name("rgb_basis.h").write_text("\n".join(out))


if __name__ == "__main__":
    main()
