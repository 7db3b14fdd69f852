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
    # the RGB path of a spectral scene still renders
    st = cr.PathTraceIntegrator(cr.Configs()).dispatch(scene, 1)
    assert st["rays"] > 0


def _uplifted(data):
    """The same scene with its RGB values uplifted explicitly."""
    sp = np.stack([np.stack([cr.rgb_to_spectrum(m.diffuseRefl), cr.rgb_to_spectrum(m.specularRefl),
                             cr.rgb_to_spectrum(m.emission)]) for m in data.materials])
    return sp, cr.rgb_to_spectrum(data.envRadiance)


def test_rgb_scene_renders_spectrally_by_uplift(gpu, orc):
    """An RGB-authored scene (no spectra) renders spectrally: bit-identical to
    the same scene given the uplifted spectra explicitly, and close to the
    oracle's spectral render of those spectra on the same streams."""
    import dataclasses
    data, camd = scenes.cornell_box()
    data.envRadiance = (0.05, 0.1, 0.2)
    size = (48, 48)
    cfg = cr.Configs(maxDepth=8, rrThreshold=0.05, useMIS=True, rrStartDepth=3, spectral=True)
    films = []
    sp, env = _uplifted(data)
    explicit = dataclasses.replace(data, spectra=sp, envSpectrum=env)
    for d in (data, explicit):
        scene = cr.Scene(d)
        sensor = cr.Sensor(size)
        cam = scenes.make_camera(camd, sensor)
        scene.setCamera(cam)
        cr.PathTraceIntegrator(cfg).dispatch(scene, 6, first_sample=2)
        films.append((sensor.data.copy(), sensor.weight.copy()))
    assert np.array_equal(films[0][0].view(np.uint32), films[1][0].view(np.uint32))
    assert np.array_equal(films[0][1].view(np.uint32), films[1][1].view(np.uint32))
    s, w, _ = orc.Scene(explicit).render(cam, size, cfg, 2, 6)
    g = films[0][0] / films[0][1][..., None]
    o = s / w[..., None]
    close = np.abs(g - o) <= 1e-4 * np.maximum(np.abs(o), 1e-3)
    assert close.mean() > 0.9, close.mean()
    assert _relmse(g, o) < 1e-3
    # the red wall stays red in XYZ: X/Y above the green wall's
    red = g[:, 2:6].reshape(-1, 3).mean(0)
    green = g[:, -6:-2].reshape(-1, 3).mean(0)
    assert red[0] / red[1] > green[0] / green[1]
