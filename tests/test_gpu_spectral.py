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


def _film(scene, camd, size, cfg, spp, first):
    sensor = cr.Sensor(size)
    scene.setCamera(scenes.make_camera(camd, sensor))
    st = cr.PathTraceIntegrator(cfg).dispatch(scene, spp, first_sample=first)
    return sensor.data.copy(), sensor.weight.copy(), st


def _materials_scene():
    """Principled, metal, Phong and Lambertian objects under an area light and
    a constant environment (RGB values: the scene uplifts them)."""
    mats = [scenes.diffuse((0.6, 0.6, 0.6)), scenes.emitter((20.0, 20.0, 20.0)),
            scenes.principled((0.8, 0.3, 0.2), 0.3, 0.0),
            scenes.principled((0.9, 0.9, 0.5), 0.1, 1.0),
            scenes.phong((0.3, 0.5, 0.2), (0.4, 0.4, 0.4), 40.0)]
    meshes = [scenes.quad((-4, 0, 4), (8, 0, 0), (0, 0, -8), 0),
              scenes.quad((-0.5, 3, -0.5), (1, 0, 0), (0, 0, 1), 1),
              scenes.icosphere((-1.1, 0.5, 0), 0.5, 2, 2),
              scenes.icosphere((0.0, 0.5, 0), 0.5, 2, 3),
              scenes.icosphere((1.1, 0.5, 0), 0.5, 2, 4)]
    data = cr.SceneData(meshes=meshes, materials=mats, envRadiance=(0.2, 0.25, 0.3))
    data.emissiveIndex = scenes._emissive_prims(data)
    camd = dict(pos=(0, 1.5, 4.5), target=(0, 0.5, 0), up=(0, 1, 0), fov=40, size=(64, 48))
    return data, camd


@pytest.mark.parametrize("mis", [True, False])
def test_spectral_wavefront_bit_identical_to_megakernel(gpu, mis):
    """Spectral transport on the wavefront: the generate kernel draws each
    path's hero bin right after its jitter and the path carries it in its
    flags through every queue; the XYZ film and the ray count are
    bit-identical to the spectral megakernel's, with one or several samples
    per pixel in flight and the per-path finish at several depths (65 =
    every bounce as wavefront launches)."""
    ctx = cr.get_context(0)
    saved = ctx.get_tuning()
    try:
        for data, camd, size in [(*scenes.cornell_box_spectral(), (48, 48)),
                                 (*_materials_scene(), (64, 48))]:
            scene = cr.Scene(data)
            base = dict(maxDepth=8, rrThreshold=0.05, useMIS=mis, rrStartDepth=3, spectral=True)
            d0, w0, s0 = _film(scene, camd, size, cr.Configs(**base), 5, first=7)
            assert np.isfinite(d0).all() and d0[..., 1].mean() > 0.0
            npix = size[0] * size[1]
            for k, fin in ((1, 65), (3, 65), (5, 1), (2, 3)):
                ctx.set_tuning(wf_pool=k * npix, wf_finish=fin)
                d1, w1, s1 = _film(scene, camd, size,
                                   cr.Configs(**base, mode=L.MODE_WAVEFRONT), 5, first=7)
                assert np.array_equal(w0.view(np.uint32), w1.view(np.uint32)), (k, fin)
                assert np.array_equal(d0.view(np.uint32), d1.view(np.uint32)), \
                    (k, fin, np.argwhere(d0 != d1)[:5])
                assert s0["rays"] == s1["rays"], (k, fin)
    finally:
        ctx.set_tuning(**saved)


def test_spectral_mode_errors(gpu):
    data, camd = scenes.cornell_box_spectral()
    scene = cr.Scene(data)
    sensor = cr.Sensor((16, 16))
    cam = scenes.make_camera(camd, sensor)
    scene.setCamera(cam)
    # spectral transport has no counting variant
    cc = cr.Configs(spectral=True).to_c()
    cc.count_work = 1
    import ctypes as C
    st = L.crh_stats()
    with pytest.raises(cr.RuntimeError):
        L.call("crh_render", scene.handle, C.byref(cam.data), C.byref(cc), 0, 1, sensor.handle,
               C.byref(st))
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
