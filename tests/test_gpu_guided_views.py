"""Guided rendering of several camera views of one scene in each launch of
the fused guided megakernel (crh_render_guided_views), on the terrain layout
at a test size (four views around the centre, glossy Principled materials,
sun disk, sky).  With a guide that does not train, every view's film is bit
for bit the film crh_render_guided makes of that view alone: the random
streams are keyed by pixel and sample index and a row of the network's
16-row tiles does not depend on the other rows.  With training on, the
guide learns from every view's vertices.  The terrain configuration's guide
is 128 wide; both widths are covered."""
import numpy as np
import pytest

import craystal_amd as cr
from craystal_amd import scenes
from craystal_amd.guiding import Guide, GuidedIntegrator

pytestmark = pytest.mark.gpu

CFG = cr.Configs(maxDepth=24, rrThreshold=0.05, useMIS=True, rrStartDepth=3)


def _terrain():
    data, cams = scenes.terrain_scene(cells=80, n_spheres=12, subdiv=2, size=(48, 27))
    cams[1]["size"] = (40, 33)          # views of different sizes
    cams[3]["size"] = (17, 64)
    return data, cams


def _gi(data, seed, materials=0, width=64):
    P = data.concatenated()[0]
    return GuidedIntegrator(CFG, Guide(P.min(0), P.max(0), seed=seed, width=width), alpha=0.5,
                            lr=1e-2, train_batch=1024, train_steps=4, materials=materials,
                            target=0)


@pytest.mark.parametrize("materials,width", [(0, 64), (1, 64), (0, 128)])
def test_views_equal_solo_guided_renders(gpu, materials, width):
    data, cams = _terrain()
    scene = cr.Scene(data)
    gi = _gi(data, 7, materials, width)
    spp, first = 3, 11
    solo = []
    for camd in cams:
        sensor = cr.Sensor(camd["size"])
        scene.setCamera(scenes.make_camera(camd, sensor))
        gi.dispatch(scene, spp, first_sample=first, train=False)
        solo.append((sensor.data.tobytes(), sensor.weight.tobytes()))
    sensors = [cr.Sensor(c["size"]) for c in cams]
    cam_objs = [scenes.make_camera(c, s) for c, s in zip(cams, sensors)]
    ctx = cr.get_context(0)
    saved = ctx.get_tuning()
    try:
        for kw in (dict(), dict(guided_pool=1 << 10, guided_grid=7),
                   dict(guided_pool=1 << 22, guided_stack=1, guided_waves=8),
                   dict(guided_stack=2, guided_waves=4, guided_grid=3)):
            ctx.set_tuning(guided_fused=1, **kw)
            for s in sensors:
                s.clear()
            st = gi.dispatch_views(scene, cam_objs, spp, first_sample=first, train=False)
            assert st["rays"] > 0
            for s, (d, w) in zip(sensors, solo):
                assert s.data.tobytes() == d and s.weight.tobytes() == w, kw
    finally:
        ctx.set_tuning(**saved)
    img = sensors[0].createImage()
    assert np.isfinite(img).all() and img.mean() > 0.01


@pytest.mark.parametrize("width", [64, 128])
def test_views_train_the_guide(gpu, width):
    data, cams = _terrain()
    scene = cr.Scene(data)
    gi = _gi(data, 9, width=width)
    sensors = [cr.Sensor(c["size"]) for c in cams]
    cam_objs = [scenes.make_camera(c, s) for c, s in zip(cams, sensors)]
    theta0 = gi.guide.get_params()
    for k in range(6):
        st = gi.dispatch_views(scene, cam_objs, 4, first_sample=100 + 4 * k, train=True)
    info = gi.guide.info()
    assert info["steps"] > 0 and np.isfinite(info["last_loss"])
    assert not np.array_equal(gi.guide.get_params(), theta0)
    assert all(np.isfinite(s.createImage()).all() for s in sensors)
