"""Spheres (SceneView::intersect's sphere loop, Scene/Scene.cu:5-29;
SphereSOA::intersect, Scene/Sphere.cu:6-49) and diffuse textures
(TextureView::sample, Core/Texture.cu:15-35; getBSDF, Material/BSDF.cu:11-15)
on the GPU against the oracle on the same inputs and random streams.

Hit IDs and t are bit-exact.  Films: the sphere (u, v) parameterisation
goes through acos/atan2, whose last-ulp results differ between the GPU and
the host library, so textured spheres can differ by rounding; the bar is the
one of the other film tests (most values equal to 1e-4, relMSE < 1e-3)."""
import numpy as np
import pytest

import craystal_amd as cr
from craystal_amd import _lib as L
from craystal_amd import scenes

pytestmark = pytest.mark.gpu


def _relmse(img, ref, eps=1e-2):
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + eps)))


def _rays(data, camd, n):
    """The mixed primary/secondary batch plus rays leaving each sphere's
    centre (inside hits take the far root)."""
    rays = scenes.mixed_batch(data, camd, n)
    rng = np.random.default_rng(17)
    extra = np.zeros(3 * 256, L.RAY_DTYPE)
    d = rng.normal(size=(len(extra), 3))
    extra["dir"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    extra["origin"] = np.repeat([s.center for s in data.spheres], 256, axis=0)
    extra["tmax"] = np.inf
    return np.concatenate([rays, extra])


@pytest.mark.parametrize("own", [False, True])
def test_trace_hits_bit_exact_with_spheres(gpu, orc, own):
    data, camd = scenes.textured_sphere_box(own)
    rays = _rays(data, camd, 1 << 16)
    scene = cr.Scene(data)
    hits = scene.trace_closest(rays)
    ref = orc.Scene(data).trace_closest(rays)
    assert np.array_equal(hits["prim"], ref["prim"])
    assert np.array_equal(hits["t"].view(np.uint32), ref["t"].view(np.uint32))
    tri = (hits["prim"] & L.CRH_HIT_SPHERE) == 0
    assert np.array_equal(hits["u"][tri].view(np.uint32), ref["u"][tri].view(np.uint32))
    sph = (hits["prim"] != L.CRH_INVALID_PRIM) & ~tri
    assert sph.sum() > 1000
    occ = scene.trace_any(rays)
    assert np.array_equal(occ, orc.Scene(data).trace_any(rays))


@pytest.mark.parametrize("own,mis,spectral", [(False, True, False), (True, True, False),
                                              (False, False, False), (True, True, True)])
def test_film_matches_oracle_with_spheres_and_textures(gpu, orc, own, mis, spectral):
    data, camd = scenes.textured_sphere_box(own)
    size = (56, 56)
    cfg = cr.Configs(maxDepth=8, rrThreshold=0.05, useMIS=mis, rrStartDepth=3,
                     spectral=spectral)
    scene = cr.Scene(data)
    sensor = cr.Sensor(size)
    cam = scenes.make_camera(camd, sensor)
    scene.setCamera(cam)
    cr.PathTraceIntegrator(cfg).dispatch(scene, 8, first_sample=5)
    s, w, _ = orc.Scene(data).render(cam, size, cfg, 5, 8)
    assert np.array_equal(sensor.weight.view(np.uint32), w.view(np.uint32))
    g = sensor.createImage()
    o = s / w[..., None]
    assert np.isfinite(g).all() and g.mean() > 0.02
    close = np.abs(g - o) <= 1e-4 * np.maximum(np.abs(o), 1e-3)
    assert close.mean() > 0.9, close.mean()
    assert _relmse(g, o) < 1e-3


def test_textures_and_spheres_change_the_image(gpu):
    """The texture and the spheres are seen: the textured back wall is not
    the flat material's colour, and removing the spheres changes the film."""
    base, camd = scenes.textured_sphere_box(True)
    size = (48, 48)
    cfg = cr.Configs(maxDepth=4, rrThreshold=0.05, useMIS=True, rrStartDepth=3)
    imgs = []
    for d in (base, cr.SceneData(meshes=base.meshes, materials=base.materials,
                                 emissiveIndex=base.emissiveIndex,
                                 textureImages=base.textureImages)):
        scene = cr.Scene(d)
        sensor = cr.Sensor(size)
        scene.setCamera(scenes.make_camera(camd, sensor))
        cr.PathTraceIntegrator(cfg).dispatch(scene, 16)
        imgs.append(sensor.createImage())
    assert np.abs(imgs[0] - imgs[1]).mean() > 1e-3
    wall = imgs[1][12:20, 12:22]           # back wall left of the tall box: checker
    assert wall[..., 2].std() > 0.02


def test_wavefront_and_batch_bit_identical_with_spheres_and_textures(gpu):
    from craystal_amd.batch import Batch
    data, camd = scenes.textured_sphere_box(True)
    size = (40, 40)
    base = dict(maxDepth=8, rrThreshold=0.05, useMIS=True, rrStartDepth=3)
    films = []
    for mode in (L.MODE_MEGAKERNEL, L.MODE_WAVEFRONT):
        scene = cr.Scene(data)
        sensor = cr.Sensor(size)
        scene.setCamera(scenes.make_camera(camd, sensor))
        cr.PathTraceIntegrator(cr.Configs(**base, mode=mode)).dispatch(scene, 4, first_sample=3)
        films.append(sensor.data.copy())
    assert np.array_equal(films[0].view(np.uint32), films[1].view(np.uint32))
    other, ocam = scenes.textured_sphere_box(False)
    scs = [cr.Scene(data), cr.Scene(other)]
    sensors = [cr.Sensor(size), cr.Sensor((24, 32))]
    cams = [scenes.make_camera(camd, sensors[0]), scenes.make_camera(ocam, sensors[1])]
    Batch(scs).render(cams, [4, 3], cr.Configs(**base), first_sample=3)
    assert np.array_equal(sensors[0].data.view(np.uint32), films[0].view(np.uint32))
    solo = cr.Sensor((24, 32))
    sc = cr.Scene(other)
    sc.setCamera(scenes.make_camera(ocam, solo))
    cr.PathTraceIntegrator(cr.Configs(**base)).dispatch(sc, 3, first_sample=3)
    assert np.array_equal(sensors[1].data.view(np.uint32), solo.data.view(np.uint32))


def test_sphere_and_texture_errors(gpu):
    data, _ = scenes.textured_sphere_box(False)
    data.spheres = data.spheres * 20          # 60 spheres > 36 triangles, no materials
    with pytest.raises(cr.RuntimeError):
        cr.Scene(data)
    data, _ = scenes.textured_sphere_box(True)
    data.materials[4].diffuseHandle = 7
    with pytest.raises(cr.RuntimeError):
        cr.Scene(data)
    data, _ = scenes.textured_sphere_box(True)
    data.sphereMaterials = [0, 1, 99]
    with pytest.raises(cr.RuntimeError):
        cr.Scene(data)
    data, _ = scenes.textured_sphere_box(True)
    data.textureImages[1] = np.zeros((2, 2, 5), np.float32)
    with pytest.raises(cr.RuntimeError):
        cr.Scene(data)
