"""The occluded-light room at its full configuration (200 boxes and
icospheres, 171,388 triangles, 1024 x 1024, max depth 16, lit only through
a 0.5 x 0.5 m opening) on the GPU.

* What the trained guide gains (the glossy room of the configuration,
  against its committed 65,536-spp plain render): at equal samples per
  pixel, guided rendering (the fused guided megakernel, one-sample MIS with
  alpha 0.5 at Lambertian vertices) with the trained guide is closer to the
  reference than the plain megakernel and than guided rendering with the
  untrained guide.  The error is relMSE without the 0.1 % of pixels with the
  largest error: the reference's Principled BSDF divides by cos(theta_o)
  twice, so single samples at grazing views can move a whole image's
  relMSE and the image mean's batch-to-batch spread understates its error
  (measured: independent 256-spp plain estimates of the mean lie up to 4.8
  estimated standard errors from the reference).
* Unbiasedness, where a 3-sigma test is valid: the same room with every
  Principled object Lambertian (same geometry and draws), whose estimator
  has bounded weights; the image means of guided rendering with the trained
  guide and of the plain megakernel agree within 3 sigma of their batch
  means."""
from pathlib import Path

import numpy as np
import pytest

import craystal_amd as cr
from craystal_amd import experiments as X
from craystal_amd import scenes
from craystal_amd.guiding import Guide, GuidedIntegrator

pytestmark = pytest.mark.gpu

SIZE = 1024
CFG = cr.Configs(maxDepth=16, rrThreshold=0.05, useMIS=True, rrStartDepth=3)
REF = Path(__file__).resolve().parents[1] / "craystal_amd" / "data" / "room1024_plain65536.npy"


def _room(glossy):
    data, camd = scenes.occluded_room(glossy=glossy)
    assert data.triangle_count == 171_388 and len(data.meshes) >= 200
    scene = cr.Scene(data)
    sensor = cr.Sensor((SIZE, SIZE))
    scene.setCamera(scenes.make_camera(camd, sensor))
    return data, scene, sensor


def _guided(data, seed, train_scene=None):
    P = data.concatenated()[0]
    gi = GuidedIntegrator(CFG, Guide(P.min(0), P.max(0), seed=seed), alpha=0.5, lr=1e-2,
                          train_batch=4096, train_steps=4, materials=1, target=0)
    if train_scene is not None:
        for k in range(8):                       # 32 spp of training
            gi.dispatch(train_scene, 4, first_sample=(1 << 26) + 4 * k, train=True,
                        readback=False)
        assert gi.guide.info()["steps"] > 0
    return gi


def _batches(render, sensor, k, spp, first):
    """k batches of spp samples: per-batch image means (k, 3) and the image
    of all of them together."""
    means, s, w = [], 0.0, 0.0
    for b in range(k):
        sensor.clear()
        render(spp, first + b * spp)
        sensor.readbackDeviceData()
        means.append(sensor.createImage().reshape(-1, 3).mean(0))
        s = s + sensor.data
        w = w + sensor.weight
    return np.array(means), s / w[..., None]


def test_room_1024_trained_guide_gains(gpu):
    data, scene, sensor = _room(glossy=True)
    ref = np.load(REF).astype(np.float32)
    assert ref.shape == (SIZE, SIZE, 3) and np.isfinite(ref).all()
    K, B, first = 8, 8, 1 << 20
    plain = cr.PathTraceIntegrator(CFG)
    _, pimg = _batches(lambda n, f: plain.dispatch(scene, n, first_sample=f, readback=False),
                       sensor, K, B, first)
    gi = _guided(data, 3, train_scene=scene)
    _, gimg = _batches(lambda n, f: gi.dispatch(scene, n, first_sample=f, train=False,
                                                readback=False), sensor, K, B, first)
    untrained = _guided(data, 3)
    _, uimg = _batches(lambda n, f: untrained.dispatch(scene, n, first_sample=f, train=False,
                                                       readback=False), sensor, K, B, first)
    e_plain = X.relmse_trimmed(pimg, ref)
    e_trained = X.relmse_trimmed(gimg, ref)
    e_untrained = X.relmse_trimmed(uimg, ref)
    print(f"trimmed relMSE at {K * B} spp: plain {e_plain:.5f}, guided (trained) "
          f"{e_trained:.5f}, guided (untrained) {e_untrained:.5f}")
    assert e_trained < 0.9 * e_plain, (e_trained, e_plain)
    assert e_trained < e_untrained, (e_trained, e_untrained)


def test_room_1024_lambertian_guided_unbiased(gpu):
    data, scene, sensor = _room(glossy=False)
    K, B, first = 16, 4, 1 << 20
    plain = cr.PathTraceIntegrator(CFG)
    pm, _ = _batches(lambda n, f: plain.dispatch(scene, n, first_sample=f, readback=False),
                     sensor, K, B, first)
    gi = _guided(data, 5, train_scene=scene)
    gm, _ = _batches(lambda n, f: gi.dispatch(scene, n, first_sample=f, train=False,
                                              readback=False), sensor, K, B, first)
    d = np.abs(pm.mean(0) - gm.mean(0))
    tol = 3 * np.sqrt(pm.var(0, ddof=1) / K + gm.var(0, ddof=1) / K)
    print(f"image means: plain {pm.mean(0)}, guided {gm.mean(0)}, 3 sigma {tol}")
    assert np.all(d <= tol), (pm.mean(0), gm.mean(0), tol)
    assert pm.mean() > 0.01
