"""The large-scale terrain layout at test sizes, on the GPU against the
oracle.

With the reference's depth cap (`leaf if count <= 4 || depth >= 20`,
Core/BVH.cu:150-212) the 43.8M-triangle terrain gets leaves of about 40
triangles; a BVH depth cap of 9 on a 17k-triangle terrain gives the same
long leaves (57 triangle tests per ray) at a size the oracle traverses in
seconds.  Hit IDs and t are bit-exact (BVHData::intersect,
Core/BVH.cu:13-84), with one ray per thread and with lanes refilling
(persistent kernels); the four camera views render in one batch that holds
the scene once, each film equal to that view rendered alone, and one view's
film matches the oracle's."""
import numpy as np
import pytest

import craystal_amd as cr
from craystal_amd import _lib as L
from craystal_amd import scenes
from craystal_amd.batch import Batch

pytestmark = pytest.mark.gpu

CFG = cr.Configs(maxDepth=24, rrThreshold=0.05, useMIS=True, rrStartDepth=3)


def _small(size=(48, 27), depth_cap=9):
    data, cams = scenes.terrain_scene(cells=80, n_spheres=12, subdiv=2, size=size)
    data.bvhParams = (32, 4, depth_cap)
    return data, cams


def test_long_leaves_hits_bit_exact(gpu, orc):
    data, cams = _small()
    gs, os_ = cr.Scene(data), orc.Scene(data)
    st = os_.bvh_stats()
    assert gs.bvh_stats()["node_count"] == st["node_count"]
    n = 1 << 16
    rays = scenes.mixed_batch(data, cams[0], n)
    oh, nv, nt = os_.trace_closest(rays, counts=True)
    assert nt.sum() / n > 10            # leaves of tens of triangles
    ctx = cr.get_context(0)
    saved = ctx.get_tuning()
    try:
        for kw in (dict(persistent=2), dict(persistent=1, refill=16, pt_grid=64),
                   dict(persistent=1, refill=1, pt_grid=5, trace_tris=1),
                   dict(persistent=1, refill=16, pt_grid=64, trace_tris=2),
                   dict(persistent=1, refill=8, pt_grid=7, trace_tris=4)):
            ctx.set_tuning(**kw)
            h = gs.trace_closest(rays)
            for k in ("prim", "t", "u", "v"):
                assert np.array_equal(h[k].view(np.uint32), oh[k].view(np.uint32)), (kw, k)
            assert np.array_equal(gs.trace_any(rays), os_.trace_any(rays)), kw
    finally:
        ctx.set_tuning(**saved)
    rb, hb = cr.DeviceBuffer(rays.nbytes), cr.DeviceBuffer(n * 16)
    rb.copyFromHost(rays)
    cnt = cr.trace_closest_device(gs, rb, n, hb, count_work=True)
    assert cnt["node_visits"] == int(nv.sum(dtype=np.uint64))
    assert cnt["tri_tests"] == int(nt.sum(dtype=np.uint64))
    assert (oh["prim"] != L.CRH_INVALID_PRIM).mean() > 0.3


def test_views_batched_share_the_scene(gpu):
    data, cams = _small()
    scene = cr.Scene(data)
    first, spp = 9, [2, 1, 3, 2]
    solo = []
    for camd, k in zip(cams, spp):
        sensor = cr.Sensor(camd["size"])
        scene.setCamera(scenes.make_camera(camd, sensor))
        cr.PathTraceIntegrator(CFG).dispatch(scene, k, first_sample=first)
        solo.append((sensor.data.copy(), sensor.weight.copy()))
    one = Batch([scene]).info()
    batch = Batch([scene] * 4)
    info = batch.info()
    assert info["scenes"] == 4 and info["triangles"] == scene.triangle_count
    assert info["pool_bytes"] == one["pool_bytes"]
    sensors = [cr.Sensor(c["size"]) for c in cams]
    cam_objs = [scenes.make_camera(c, s) for c, s in zip(cams, sensors)]
    for rep in range(2):                # capture, then replay
        for s in sensors:
            s.clear()
        batch.render(cam_objs, spp, CFG, first_sample=first)
        for s, (d, w) in zip(sensors, solo):
            assert s.data.tobytes() == d.tobytes() and s.weight.tobytes() == w.tobytes()
    assert np.isfinite(solo[0][0]).all() and solo[0][0].mean() > 0.01


def test_view_film_matches_oracle(gpu, orc):
    data, cams = _small(size=(40, 24))
    camd = cams[1]
    size = camd["size"]
    scene = cr.Scene(data)
    sensor = cr.Sensor(size)
    cam = scenes.make_camera(camd, sensor)
    scene.setCamera(cam)
    cr.PathTraceIntegrator(CFG).dispatch(scene, 4, first_sample=3)
    s, w, _ = orc.Scene(data).render(cam, size, CFG, 3, 4)
    assert np.array_equal(sensor.weight.view(np.uint32), w.view(np.uint32))
    g = sensor.createImage()
    o = s / w[..., None]
    close = np.abs(g - o) <= 1e-4 * np.maximum(np.abs(o), 1e-3)
    assert close.mean() > 0.9, close.mean()
    assert float(np.mean((g - o) ** 2 / (o ** 2 + 1e-2))) < 1e-3


def test_fast_bvh_parameters_bit_exact(gpu, orc):
    """The terrain's fast parameter set (no effective depth cap), built on
    the GPU, on a 256k-triangle terrain: node arrays and primitive order
    equal the oracle's build with the same parameters, and hit IDs, t and
    the node / triangle counts equal the oracle's traversal of that tree;
    leaves keep to the 4-triangle rule where the reference cap gives long
    ones."""
    data, cams = scenes.terrain_scene(cells=320, n_spheres=40, subdiv=3, size=(48, 27))
    data.bvhParams = scenes.TERRAIN_FAST_BVH
    data.bvhDevice = True
    gs, os_ = cr.Scene(data), orc.Scene(data)
    gst, ost = gs.bvh_stats(), os_.bvh_stats()
    for k in ("node_count", "max_depth", "max_leaf", "root"):
        assert gst[k] == ost[k], k
    assert gst["max_leaf"] <= 12 and gst["max_depth"] < 60
    gn, gp = gs.export_bvh()
    on, op = os_.export_bvh()
    assert gn.tobytes() == on.tobytes() and np.array_equal(gp, op)
    n = 1 << 16
    rays = scenes.mixed_batch(data, cams[0], n)
    oh, nv, nt = os_.trace_closest(rays, counts=True)
    assert nt.sum() / n < 10
    ctx = cr.get_context(0)
    saved = ctx.get_tuning()
    try:
        for kw in (dict(persistent=2, stack=3), dict(persistent=1, refill=16, pt_grid=64),
                   dict(persistent=1, refill=32, pt_grid=4096, trace_tris=2, stack=2)):
            ctx.set_tuning(**kw)
            h = gs.trace_closest(rays)
            for k in ("prim", "t", "u", "v"):
                assert np.array_equal(h[k].view(np.uint32), oh[k].view(np.uint32)), (kw, k)
    finally:
        ctx.set_tuning(**saved)
    rb, hb = cr.DeviceBuffer(rays.nbytes), cr.DeviceBuffer(n * 16)
    rb.copyFromHost(rays)
    cnt = cr.trace_closest_device(gs, rb, n, hb, count_work=True)
    assert cnt["node_visits"] == int(nv.sum(dtype=np.uint64))
    assert cnt["tri_tests"] == int(nt.sum(dtype=np.uint64))
