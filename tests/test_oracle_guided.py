"""Pins of the oracle's guided estimator (an addition with no reference
counterpart): one-sample MIS of the reference's BSDF sampling (fraction
alpha) and the guide's vMF mixture, weighted f * p_act / (q_mix * p_rep) so
that its expectation is the plain estimator's (Core/Integrator.cu:195-203
with Material/BSDF.cu:51-103 sampling).

What fixes the answers independently of the code under test:
* p_act, the density of BSDF::sampleLocal's directions, against histograms
  of those directions over the whole sphere (it must also integrate to 1);
* the mixture's head, density and sampler against the float64 model of
  oracle/guiding.py (pinned by quadrature and histograms in
  tests/test_guiding_oracle.py);
* a proposal equal to the BSDF's own procedure gives, for every direction,
  the plain estimator's weight f / p_rep (with the guide equal to the BSDF,
  the guided estimator reduces to the plain path);
* guided images: the diffuse plane under a constant environment equals the
  closed form 0.95 rho E, and on a Principled scene the guided mean equals
  the plain oracle's within 3 sigma.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import craystal_amd as cr
from craystal_amd import scenes
from oracle import guiding as G

MATERIALS = {
    "lambert": scenes.diffuse((0.6, 0.4, 0.2)),
    "principled-r0.4-m0": scenes.principled((0.7, 0.5, 0.3), 0.4, 0.0),
    "principled-r0.6-m1": scenes.principled((0.7, 0.5, 0.3), 0.6, 1.0),
    "principled-r0.3-m0.5": scenes.principled((0.2, 0.8, 0.5), 0.3, 0.5),
}


def _dir(theta_deg, phi_deg=20.0):
    t, p = math.radians(theta_deg), math.radians(phi_deg)
    return np.array([math.sin(t) * math.cos(p), math.sin(t) * math.sin(p), math.cos(t)],
                    np.float32)


def _outputs(seed, scale=1.5):
    """48 guide outputs with distinct lobes (kappa logits spread over 0..4)."""
    rng = np.random.default_rng(seed)
    o = rng.normal(size=48) * scale
    for j in range(8):
        o[5 * j + 4] = rng.uniform(-1.0, 4.0)
    return o.astype(np.float32)


@pytest.mark.parametrize("name", list(MATERIALS))
@pytest.mark.parametrize("theta", [0.0, 40.0, 72.0])
def test_actual_pdf_is_the_density_of_the_bsdf_sampler(orc, name, theta):
    mat = MATERIALS[name]
    wo = _dir(theta)
    n = 400_000
    u = np.random.default_rng(17).random((n, 2)).astype(np.float32)
    wi = orc.bsdf_sample_dirs(mat, wo, u).astype(np.float64)
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    nz, nphi = 24, 24                        # equal-area cells in (z, phi)
    zi = np.clip(((wi[:, 2] + 1) / 2 * nz).astype(int), 0, nz - 1)
    pi_ = np.clip(((np.arctan2(wi[:, 1], wi[:, 0]) % (2 * np.pi)) / (2 * np.pi) * nphi)
                  .astype(int), 0, nphi - 1)
    obs = np.bincount(zi * nphi + pi_, minlength=nz * nphi).astype(np.float64)
    # expected counts: n * integral of p_act over each cell (48 x 8 midpoint
    # grid: the glossy lobes are narrow in z near the pole)
    kz, kp = 48, 8
    zz = -1 + (np.arange(nz * kz) + 0.5) * (2 / (nz * kz))
    pp = (np.arange(nphi * kp) + 0.5) * (2 * np.pi / (nphi * kp))
    Z, Pp = np.meshgrid(zz, pp, indexing="ij")
    r = np.sqrt(np.maximum(0, 1 - Z * Z))
    d = np.stack([r * np.cos(Pp), r * np.sin(Pp), Z], -1).reshape(-1, 3)
    dens = orc.actual_pdf_batch(mat, wo, d).astype(np.float64).reshape(nz * kz, nphi * kp)
    cell = (2 / (nz * kz)) * (2 * np.pi / (nphi * kp))
    expct = (dens * cell).reshape(nz, kz, nphi, kp).sum((1, 3)).reshape(-1) * n
    assert abs(expct.sum() / n - 1.0) < 0.01, expct.sum() / n     # a density: integrates to 1
    big = expct > 50
    z = (obs[big] - expct[big]) / np.sqrt(expct[big])
    assert np.abs(z).max() < 6.0 and np.mean(z * z) < 2.0, (np.abs(z).max(), np.mean(z * z))
    assert obs[~big].sum() < 0.02 * n + 5 * math.sqrt(max(expct[~big].sum(), 1))


def test_mixture_matches_the_float64_model(orc):
    rng = np.random.default_rng(3)
    for seed in range(6):
        o = _outputs(seed)
        w, mu, kap = orc.mixture(o)
        gw, gmu, gk = G.head(o.astype(np.float64)[None, :])
        assert np.allclose(w, gw[0], atol=1e-6) and np.allclose(mu, gmu[0], atol=1e-6)
        assert np.allclose(kap, gk[0], rtol=1e-5)
        for _ in range(40):
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            ref = G.mixture_pdf(o.astype(np.float64)[None, :], d[None, :])[0]
            assert math.isclose(orc.mixture_pdf(o, d), ref, rel_tol=2e-4, abs_tol=1e-7)
            u = rng.random(3)
            cdf = np.cumsum(gw[0])
            if np.min(np.abs(cdf - u[0])) < 1e-5:
                continue
            j = min(int(np.searchsorted(cdf, u[0], side="right")), 7)
            ref_d = G.sample_vmf(gmu[0, j], np.array(gk[0, j]), np.array(u[1]), np.array(u[2]))
            got = orc.mixture_sample(o, u.astype(np.float32))
            assert np.allclose(got, ref_d / np.linalg.norm(ref_d), atol=2e-4), (got, ref_d)


@pytest.mark.parametrize("name", list(MATERIALS))
def test_guide_equal_to_the_bsdf_gives_the_plain_weight(orc, name):
    """With the BSDF's own procedure as the proposal, q_mix = p_act, so the
    guided weight f / pdf is the plain weight f / p_rep at every direction,
    whichever technique drew it."""
    mat = MATERIALS[name]
    rng = np.random.default_rng(5)
    checked = 0
    for _ in range(400):
        wo = rng.normal(size=3)
        wo[2] = abs(wo[2]) + 0.05
        wo = (wo / np.linalg.norm(wo)).astype(np.float32)
        u4 = rng.random(4).astype(np.float32)
        wi, f, pdf = orc.guided_sample_local(mat, None, 0.5, wo, u4)
        f2, prep = orc.bsdf_eval_local(mat, wo, wi)
        assert np.array_equal(f, f2)
        if pdf > 0 and prep > 0 and f.max() > 0:
            assert np.allclose(f / pdf, f2 / prep, rtol=2e-6), (f / pdf, f2 / prep)
            checked += 1
        # the same uniforms through the BSDF branch draw the plain sample
        if u4[0] < 0.5:
            wi_p, _, _ = orc.bsdf_sample_local(mat, wo, u4[1:3])
            assert np.array_equal(wi, wi_p)
    assert checked > 250


def test_guided_weight_of_a_mixture_direction(orc):
    """A mixture-drawn direction: weight = f * p_act / ((a p_act + (1-a) q) p_rep),
    with q the float64 mixture density there."""
    mat = MATERIALS["principled-r0.4-m0"]
    o = _outputs(11)
    rng = np.random.default_rng(9)
    n_ok = 0
    for _ in range(200):
        wo = _dir(rng.uniform(0, 80), rng.uniform(0, 360))
        u4 = np.concatenate([[0.75], rng.random(3)]).astype(np.float32)   # mixture branch
        wi, f, pdf = orc.guided_sample_local(mat, o, 0.5, wo, u4)
        f2, prep = orc.bsdf_eval_local(mat, wo, wi)
        pact = orc.actual_pdf_local(mat, wo, wi)
        q = G.mixture_pdf(o.astype(np.float64)[None, :], wi.astype(np.float64)[None, :])[0]
        if prep > 0 and pact > 0:
            expect = (0.5 * pact + 0.5 * q) * prep / pact
            assert math.isclose(pdf, expect, rel_tol=1e-4), (pdf, expect)
            n_ok += 1
        else:
            assert pdf == 0.0
    assert n_ok > 50


def _cam(pos, target, fov, size):
    return scenes.make_camera(dict(pos=pos, target=target, up=(0, 0, -1), fov=fov, size=size))


@pytest.mark.parametrize("bsdf_as_guide", [False, True])
def test_guided_plane_under_environment_closed_form(orc, bsdf_as_guide):
    """Diffuse plane (albedo rho) under a constant environment E: the plain
    estimator gives 0.95 rho E (one bounce, the unconditional 5 % kill); the
    guided estimator, with any proposal, must give the same mean."""
    rho, E = 0.5, 2.0
    data = cr.SceneData(meshes=[scenes.quad((-50, 0, 50), (100, 0, 0), (0, 0, -100), 0)],
                        materials=[scenes.diffuse((rho,) * 3)], envRadiance=(E,) * 3)
    cam = _cam((0, 5, 0), (0, 0, 0), 30, (64, 64))
    s = orc.Scene(data)
    o = _outputs(21)
    cfg = cr.Configs(maxDepth=10, rrThreshold=0.05)
    vals = np.concatenate([
        s.render_samples_guided(cam, (64, 64), cfg, None if bsdf_as_guide else o, k,
                                bsdf_as_guide=bsdf_as_guide)[0][..., 0].ravel()
        for k in range(12)])
    se = vals.std() / math.sqrt(len(vals))
    assert abs(vals.mean() - 0.95 * rho * E) < 4 * se + 1e-4, (vals.mean(), 0.95 * rho * E, se)
    if bsdf_as_guide:   # every weight is rho, as in the plain estimator
        assert set(np.unique(np.round(vals, 4))) <= {0.0, round(rho * E, 4)}
    else:
        assert vals.std() > 0.3                  # the weights really vary


def test_guided_cornell_means_equal_plain(orc):
    """Cornell box with Principled boxes, MIS on: per-channel image means of
    the guided oracle (a fixed mixture, and the BSDF as the proposal) within
    3 sigma of the plain oracle's, over independent 1-spp passes."""
    data, camd = scenes.cornell_box()
    data.materials[0] = scenes.principled((0.7, 0.6, 0.5), 0.35, 0.6)
    s = orc.Scene(data)
    size = (32, 32)
    cam = scenes.make_camera(camd)
    cfg = cr.Configs(maxDepth=8, rrThreshold=0.05, useMIS=True, rrStartDepth=3)
    o = _outputs(4)
    K = 32
    plain = np.array([s.render_samples(cam, size, cfg, k)[0].reshape(-1, 3).mean(0)
                      for k in range(K)])
    for out48, as_bsdf in ((o, False), (None, True)):
        gm = np.array([s.render_samples_guided(cam, size, cfg, out48, 5000 + k,
                                               bsdf_as_guide=as_bsdf)[0].reshape(-1, 3).mean(0)
                       for k in range(K)])
        for ch in range(3):
            se = math.sqrt(plain[:, ch].var(ddof=1) / K + gm[:, ch].var(ddof=1) / K)
            assert abs(plain[:, ch].mean() - gm[:, ch].mean()) <= 3 * se, \
                (as_bsdf, ch, plain[:, ch].mean(), gm[:, ch].mean(), se)
