"""GPU megakernel vs the oracle's restatement of pathTraceKernel
(Core/Integrator.cu:142-227) on the same seeded (pixel, sample) streams.

Both sides draw identical random numbers (Sampler keyed by pixel and sample
index) and share every + - * / sqrt bit; they differ only where single
precision transcendentals (sin, cos, pow, acos) differ in the last ulp, which
can flip a rare branch.  The bar: image means per channel within 3 sigma of
the per-pass spread, relMSE (eps 1e-2) small, and the large majority of
pixels equal to ~1e-4 relative.
"""
import numpy as np
import pytest

import craystal_amd as cr
from craystal_amd import scenes

pytestmark = pytest.mark.gpu


def _render_pair(data, camd, size, cfg, spp, orc, first=0):
    scene = cr.Scene(data)
    sensor = cr.Sensor(size)
    cam = scenes.make_camera(camd, sensor)
    scene.setCamera(cam)
    cr.PathTraceIntegrator(cfg).dispatch(scene, spp, first_sample=first)
    osc = orc.Scene(data)
    s, w, _ = osc.render(cam, size, cfg, first, spp)
    return sensor, s, w, scene, cam, osc


def _relmse(img, ref, eps=1e-2):
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + eps)))


@pytest.mark.parametrize("mis", [True, False])
def test_cornell_film_matches_oracle(gpu, orc, mis):
    data, camd = scenes.cornell_box()
    size = (64, 64)
    cfg = cr.Configs(maxDepth=8, rrThreshold=0.05, useMIS=mis, rrStartDepth=3)
    sensor, s, w, *_ = _render_pair(data, camd, size, cfg, 8, orc)
    # weights depend only on the jitter: identical on both sides
    assert np.array_equal(sensor.weight.view(np.uint32), w.view(np.uint32))
    g = sensor.createImage()
    o = s / w[..., None]
    assert np.isfinite(g).all()
    close = np.abs(g - o) <= 1e-4 * np.maximum(np.abs(o), 1e-3)
    assert close.mean() > 0.9, f"only {close.mean():.3f} of values match to 1e-4"
    assert _relmse(g, o) < 1e-3


def test_cornell_means_within_3_sigma(gpu, orc):
    """Per-channel image means over 16 independent 1-spp passes: the GPU mean
    lies within 3 sigma (standard error of the oracle's per-pass means)."""
    data, camd = scenes.cornell_box()
    size = (48, 48)
    cfg = cr.Configs(maxDepth=8, rrThreshold=0.05, useMIS=True, rrStartDepth=3)
    scene = cr.Scene(data)
    osc = orc.Scene(data)
    gm, om = [], []
    for k in range(16):
        sensor = cr.Sensor(size)
        cam = scenes.make_camera(camd, sensor)
        scene.setCamera(cam)
        cr.PathTraceIntegrator(cfg).dispatch(scene, 1, first_sample=k)
        gm.append(sensor.createImage().reshape(-1, 3).mean(0))
        Lr, _, _ = osc.render_samples(cam, size, cfg, 1000 + k)
        om.append(Lr.reshape(-1, 3).mean(0))
    gm, om = np.array(gm), np.array(om)
    se = np.sqrt(gm.var(0, ddof=1) / len(gm) + om.var(0, ddof=1) / len(om))
    assert (np.abs(gm.mean(0) - om.mean(0)) <= 3 * se + 1e-7).all(), (gm.mean(0), om.mean(0), se)


def test_principled_phong_scene(gpu, orc):
    """Spheres with Principled (metal/dielectric) and Phong materials under an
    area light plus constant environment."""
    mats = [scenes.diffuse((0.6, 0.6, 0.6)), scenes.emitter((20.0, 20.0, 20.0)),
            scenes.principled((0.8, 0.3, 0.2), 0.3, 0.0), scenes.principled((0.9, 0.9, 0.5), 0.1, 1.0),
            scenes.phong((0.3, 0.5, 0.2), (0.4, 0.4, 0.4), 40.0)]
    meshes = [scenes.quad((-4, 0, 4), (8, 0, 0), (0, 0, -8), 0),
              scenes.quad((-0.5, 3, -0.5), (1, 0, 0), (0, 0, 1), 1),
              scenes.icosphere((-1.1, 0.5, 0), 0.5, 2, 2),
              scenes.icosphere((0.0, 0.5, 0), 0.5, 2, 3),
              scenes.icosphere((1.1, 0.5, 0), 0.5, 2, 4)]
    data = cr.SceneData(meshes=meshes, materials=mats, envRadiance=(0.2, 0.25, 0.3))
    data.emissiveIndex = scenes._emissive_prims(data)
    camd = dict(pos=(0, 1.5, 4.5), target=(0, 0.5, 0), up=(0, 1, 0), fov=40, size=(64, 48))
    cfg = cr.Configs(maxDepth=8, rrThreshold=0.05, useMIS=True, rrStartDepth=3)
    sensor, s, w, *_ = _render_pair(data, camd, (64, 48), cfg, 8, orc)
    g = sensor.createImage()
    o = s / w[..., None]
    assert np.isfinite(g).all()
    close = np.abs(g - o) <= 1e-3 * np.maximum(np.abs(o), 1e-3)
    assert close.mean() > 0.8, close.mean()
    assert _relmse(g, o) < 5e-3


def test_render_deterministic_and_progressive(gpu):
    """Two renders are bit-identical (gather film, no atomics), and rendering
    [0,4) then [4,8) into one film equals rendering [0,8) at once."""
    data, camd = scenes.cornell_box()
    size = (40, 40)
    cfg = cr.Configs(maxDepth=8, rrThreshold=0.05, useMIS=True, rrStartDepth=3)
    scene = cr.Scene(data)
    outs = []
    for split in (False, False, True):
        sensor = cr.Sensor(size)
        cam = scenes.make_camera(camd, sensor)
        scene.setCamera(cam)
        it = cr.PathTraceIntegrator(cfg)
        if split:
            it.dispatch(scene, 4, first_sample=0)
            it.dispatch(scene, 4, first_sample=4)
        else:
            it.dispatch(scene, 8, first_sample=0)
        outs.append((sensor.data.copy(), sensor.weight.copy()))
    for d, w in outs[1:]:
        assert np.array_equal(d.view(np.uint32), outs[0][0].view(np.uint32))
        assert np.array_equal(w.view(np.uint32), outs[0][1].view(np.uint32))


def test_work_counters_match_oracle_trace(gpu, orc):
    """The counting megakernel's ray count equals the oracle's ray count for
    the same pass (same RNG streams -> same number of traversals), less the
    MIS BSDF rays that cannot hit their sampled emitter triangle, which the
    GPU does not trace (the film is the same either way)."""
    import ctypes as C
    from craystal_amd import _lib as L
    data, camd = scenes.cornell_box()
    size = (32, 32)
    cfg = cr.Configs(maxDepth=8, rrThreshold=0.05, useMIS=True, rrStartDepth=3)
    scene = cr.Scene(data)
    sensor = cr.Sensor(size)
    cam = scenes.make_camera(camd, sensor)
    c = cfg.to_c()
    c.count_work = 1
    st = L.crh_stats()
    L.call("crh_render", scene.handle, C.byref(cam.data), C.byref(c), 3, 1, sensor.handle,
           C.byref(st))
    _, _, orays, unreach = orc.Scene(data).render_samples(cam, size, cfg, 3, counts=True)
    assert 0 < unreach < orays
    traced = int(orays) - int(unreach)
    assert abs(int(st.rays) - traced) <= max(2, traced // 500)
    assert st.node_visits > 0 and st.tri_tests > 0


def _film(scene, camd, size, cfg, spp, first=0):
    sensor = cr.Sensor(size)
    cam = scenes.make_camera(camd, sensor)
    scene.setCamera(cam)
    st = cr.PathTraceIntegrator(cfg).dispatch(scene, spp, first_sample=first)
    return sensor.data.copy(), sensor.weight.copy(), st


@pytest.mark.parametrize("mis", [True, False])
def test_wavefront_film_bit_identical_to_megakernel(gpu, mis):
    """Same RNG streams, same per-path arithmetic (integrator.h), different
    schedule: the two films are bit-identical, and so is the ray count, with
    one sample per pixel in flight or several (path pools of 1, 3 and 5
    samples per pixel over 5 passes, the last group ragged)."""
    from craystal_amd import _lib as L
    ctx = cr.get_context(0)
    saved = ctx.get_tuning()
    try:
        for data, camd, size in [(*scenes.cornell_box(), (64, 64)),
                                 (*_materials_scene(), (64, 48))]:
            scene = cr.Scene(data)
            base = dict(maxDepth=8, rrThreshold=0.05, useMIS=mis, rrStartDepth=3)
            d0, w0, s0 = _film(scene, camd, size, cr.Configs(**base), 5, first=5)
            npix = size[0] * size[1]
            # (samples per pixel in flight, bounces before the per-path finish; 65 = none)
            for k, fin in ((1, 65), (3, 65), (5, 65), (3, 1), (1, 2), (5, 5), (2, 8)):
                ctx.set_tuning(wf_pool=k * npix, wf_finish=fin)
                d1, w1, s1 = _film(scene, camd, size,
                                   cr.Configs(**base, mode=L.MODE_WAVEFRONT), 5, first=5)
                assert np.array_equal(w0.view(np.uint32), w1.view(np.uint32)), (k, fin)
                assert np.array_equal(d0.view(np.uint32), d1.view(np.uint32)), \
                    (k, fin, np.argwhere(d0 != d1)[:5])
                assert s0["rays"] == s1["rays"], (k, fin)
    finally:
        ctx.set_tuning(**saved)


@pytest.mark.parametrize("spectral", [False, True])
def test_path_regeneration_bit_identical(gpu, spectral):
    """Persistent megakernel whose lanes start a new path as soon as theirs
    ends vs one path per thread: bit-identical films and ray counts, over
    image sizes that are not multiples of the 16x16 tile, pass groups of 1
    and 3 samples per pixel (the last group ragged), MIS on and off, depth
    limits 0, 1 and 8 and persistent grids smaller and larger than the
    work."""
    from craystal_amd import _lib as L
    ctx = cr.get_context(0)
    saved = ctx.get_tuning()
    if spectral:
        cases = [(*scenes.cornell_box_spectral(), (37, 29))]
    else:
        cases = [(*scenes.cornell_box(), (64, 64)), (*_materials_scene(), (61, 45))]
    try:
        for data, camd, size in cases:
            scene = cr.Scene(data)
            npix = size[0] * size[1]
            for mis, depth in [(True, 8), (False, 8), (True, 1), (True, 0)]:
                cfg = cr.Configs(maxDepth=depth, rrThreshold=0.05, useMIS=mis, rrStartDepth=3,
                                 spectral=spectral)
                ctx.set_tuning(render_regen=2, render_pool=npix)
                d0, w0, s0 = _film(scene, camd, size, cfg, 4, first=7)
                for pool, grid in [(npix, 1), (npix, 3), (3 * npix, 2048)]:
                    ctx.set_tuning(render_regen=1, render_pool=pool, render_grid=grid)
                    d1, w1, s1 = _film(scene, camd, size, cfg, 4, first=7)
                    key = (size, mis, depth, pool // npix, grid)
                    assert np.array_equal(w0.view(np.uint32), w1.view(np.uint32)), key
                    assert np.array_equal(d0.view(np.uint32), d1.view(np.uint32)), \
                        (key, np.argwhere(d0 != d1)[:5])
                    assert s0["rays"] == s1["rays"], key
                if depth:
                    assert d0.max() > 0
    finally:
        ctx.set_tuning(**saved)


def _materials_scene():
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


def test_wavefront_large_image_matches_megakernel(gpu):
    """1080p passes on the 983k-triangle scene (queues of 2M paths and more),
    every bounce as wavefront launches and with the per-path finish after
    1 and 3 bounces (at this size the threads finishing a path's extension
    entry and its MIS entry run far apart, which the small images do not
    exercise)."""
    from craystal_amd import _lib as L
    data, camd = scenes.spheres_scene()
    scene = cr.Scene(data)
    base = dict(maxDepth=16, rrThreshold=0.05, useMIS=True, rrStartDepth=3)
    ctx = cr.get_context(0)
    saved = ctx.get_tuning()
    try:
        d0, w0, s0 = _film(scene, camd, (1920, 1080), cr.Configs(**base), 2, first=2)
        for fin, pool in ((65, 1 << 21), (1, 1 << 22), (3, 1 << 21)):
            ctx.set_tuning(wf_finish=fin, wf_pool=pool)
            d1, w1, s1 = _film(scene, camd, (1920, 1080),
                               cr.Configs(**base, mode=L.MODE_WAVEFRONT), 2, first=2)
            assert np.array_equal(d0.view(np.uint32), d1.view(np.uint32)), \
                (fin, int((d0 != d1).any(-1).sum()))
            assert s0["rays"] == s1["rays"] > 4 * 1920 * 1080, fin
    finally:
        ctx.set_tuning(**saved)
