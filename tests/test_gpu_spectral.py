"""Spectral transport on the GPU (three of the 95 wavelength bins per path,
CIE XYZ film) against the oracle's restatement on the same seeded streams."""
import numpy as np
import pytest

import craystal_amd as cr
from craystal_amd import _lib as L
from craystal_amd import scenes

pytestmark = pytest.mark.gpu


def _relmse(img, ref, eps=1e-2):
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + eps)))


@pytest.mark.parametrize("mis", [True, False])
def test_spectral_cornell_matches_oracle(gpu, orc, mis):
    data, camd = scenes.cornell_box_spectral()
    size = (48, 48)
    cfg = cr.Configs(maxDepth=8, rrThreshold=0.05, useMIS=mis, rrStartDepth=3, spectral=True)
    scene = cr.Scene(data)
    sensor = cr.Sensor(size)
    cam = scenes.make_camera(camd, sensor)
    scene.setCamera(cam)
    cr.PathTraceIntegrator(cfg).dispatch(scene, 8, first_sample=3)
    s, w, _ = orc.Scene(data).render(cam, size, cfg, 3, 8)
    assert np.array_equal(sensor.weight.view(np.uint32), w.view(np.uint32))
    g = sensor.createImage()
    o = s / w[..., None]
    assert np.isfinite(g).all() and g[..., 1].mean() > 0.01
    close = np.abs(g - o) <= 1e-4 * np.maximum(np.abs(o), 1e-3)
    assert close.mean() > 0.9, close.mean()
    assert _relmse(g, o) < 1e-3
    # the red wall reflects long wavelengths: X/Y above the green wall's
    red = g[:, 2:6].reshape(-1, 3).mean(0)
    green = g[:, -6:-2].reshape(-1, 3).mean(0)
    assert red[0] / red[1] > green[0] / green[1]


def test_spectral_mode_errors(gpu):
    data, camd = scenes.cornell_box_spectral()
    scene = cr.Scene(data)
    sensor = cr.Sensor((16, 16))
    scene.setCamera(scenes.make_camera(camd, sensor))
    with pytest.raises(cr.RuntimeError):
        cr.PathTraceIntegrator(cr.Configs(spectral=True, mode=L.MODE_WAVEFRONT)).dispatch(scene, 1)
    plain, _ = scenes.cornell_box()
    p = cr.Scene(plain)
    p.setCamera(scenes.make_camera(camd, cr.Sensor((16, 16))))
    with pytest.raises(cr.RuntimeError):
        cr.PathTraceIntegrator(cr.Configs(spectral=True)).dispatch(p, 1)
    # the RGB path of a spectral scene still renders
    st = cr.PathTraceIntegrator(cr.Configs()).dispatch(scene, 1)
    assert st["rays"] > 0
