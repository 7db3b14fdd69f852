"""RGB -> 95-bin spectrum uplift (crh_rgb_to_spectrum), the replacement of
the reference's unimplemented BroadSpectrum::fromRGB (Core/Spectrum.cpp:49).
Host code, CPU only.

What fixes the answers independently of the product: the colour of the
result, computed in float64 from the oracle's colour matching functions and
the sRGB matrix, must be the input (Smits's construction makes every basis
spectrum round-trip exactly and the decomposition is linear in the sorted
channel differences, so any wrong basis row, weight or branch breaks it);
grays are flat; reflectances stay in [0, 1]; the map is homogeneous and
continuous across the six orderings of the channels.
"""
import itertools

import numpy as np
import pytest

import craystal_amd as cr
from craystal_amd.core import SRGB_TO_XYZ


@pytest.fixture(scope="module")
def rgb_e(orc):
    """(3, 95) float64: spectrum -> linear sRGB white-balanced so that the
    flat spectrum 1 is (1, 1, 1), from the oracle's CIE table."""
    c = np.zeros((95, 3), np.float32)
    orc.lib().orc_cie_cmf(c.ctypes.data)
    c = c.astype(np.float64).T
    to_xyz = c / c[1].sum()
    m = np.linalg.inv(SRGB_TO_XYZ.astype(np.float64)) @ to_xyz
    return m / m.sum(axis=1)[:, None]


def _up(rgb):
    return cr.rgb_to_spectrum(np.asarray(rgb, np.float32)).astype(np.float64)


def test_basis_colours_round_trip(rgb_e):
    for rgb in itertools.product((0.0, 1.0), repeat=3):
        s = _up(rgb)
        assert s.min() >= 0.0 and s.max() <= 1.0 + 1e-6, rgb
        assert np.allclose(rgb_e @ s, rgb, atol=2e-5), (rgb, rgb_e @ s)


def test_random_colours_round_trip_and_bounds(rgb_e):
    rng = np.random.default_rng(11)
    for rgb in rng.random((2000, 3)):
        s = _up(rgb)
        assert s.min() >= -1e-7 and s.max() <= 1.0 + 1e-6
        assert np.allclose(rgb_e @ s, rgb, atol=3e-5), rgb
    for rgb in rng.random((200, 3)) * 60.0:                     # emission
        assert np.allclose(rgb_e @ _up(rgb), rgb, rtol=2e-5, atol=1e-4)


def test_grays_are_flat_and_negative_channels_clamp():
    for k in (0.0, 0.25, 0.73, 1.0, 12.0):
        assert np.array_equal(cr.rgb_to_spectrum((k, k, k)), np.full(95, k, np.float32))
    assert np.array_equal(cr.rgb_to_spectrum((-1.0, 0.5, 0.5)),
                          cr.rgb_to_spectrum((0.0, 0.5, 0.5)))


def test_homogeneous_and_continuous_across_orderings():
    rng = np.random.default_rng(3)
    for rgb in rng.random((300, 3)):
        assert np.allclose(_up(3.5 * rgb), 3.5 * _up(rgb), rtol=1e-6, atol=1e-6)
    eps = 1e-4
    for a, b in rng.random((200, 2)):
        # two equal channels: the neighbouring orderings meet without a jump
        for base in ((a, a, b), (a, b, a), (b, a, a)):
            s0 = _up(base)
            for d in itertools.product((-eps, eps), repeat=3):
                assert np.abs(_up(np.add(base, d)) - s0).max() < 20 * eps


def test_white_point_is_the_flat_spectrum():
    w = cr.spectrum_to_xyz(cr.rgb_to_spectrum((1.0, 1.0, 1.0)))
    assert np.allclose(w, cr.spectrum_to_xyz(np.ones(95, np.float32)))
    # Y of the flat 1 under the reference's normalisation: kXYZStep (470/95)
    # over the table's spacing (469/94)
    assert abs(w[1] - (470 / 95) / (469 / 94)) < 1e-4
