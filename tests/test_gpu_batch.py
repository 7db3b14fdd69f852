"""Batched rendering of many scenes in one persistent launch per sample
index: every film equals the scene rendered alone, bit for bit, whatever the
tile shape, grid, work ranges, stack placement or graph replay."""
import numpy as np
import pytest

import craystal_amd as cr
from craystal_amd import _lib as L
from craystal_amd import autotune, scenes
from craystal_amd.batch import Batch

pytestmark = pytest.mark.gpu

CFG = cr.Configs(maxDepth=12, rrThreshold=0.05, useMIS=True, rrStartDepth=3)


def _small_batch(n=6, seed=5):
    out = []
    for i in range(n):
        data, camd = scenes.batch_scene(i, seed=seed, tri_range=(1_000, 40_000),
                                        resolutions=(24, 40, 64), spps=(1, 2, 3))
        out.append((data, camd))
    return out


def _films(sensors):
    return [(s.data.copy(), s.weight.copy()) for s in sensors]


def _solo(scene_list, camds, spp, first):
    res = []
    for sc, camd, k in zip(scene_list, camds, spp):
        sensor = cr.Sensor(camd["size"])
        cam = scenes.make_camera(camd, sensor)
        sc.setCamera(cam)
        if k:
            cr.PathTraceIntegrator(CFG).dispatch(sc, k, first_sample=first)
        else:
            sensor.readbackDeviceData()
        res.append((sensor.data.copy(), sensor.weight.copy()))
    return res


def _same(a, b):
    return all(x[0].tobytes() == y[0].tobytes() and x[1].tobytes() == y[1].tobytes()
               for x, y in zip(a, b))


def test_batch_equals_solo_renders(gpu):
    items = _small_batch()
    scene_list = [cr.Scene(d) for d, _ in items]
    camds = [c for _, c in items]
    spp = [c["spp"] for c in camds]
    spp[1] = 0                          # a scene with nothing to render this call
    first = 77
    solo = _solo(scene_list, camds, spp, first)
    batch = Batch(scene_list)
    info = batch.info()
    assert info["scenes"] == len(items) and info["triangles"] == sum(
        s.triangle_count for s in scene_list)
    del scene_list                      # the batch owns a copy of the scene data
    sensors = [cr.Sensor(c["size"]) for c in camds]
    cams = [scenes.make_camera(c, s) for c, s in zip(camds, sensors)]
    ctx = cr.get_context(0)
    saved = ctx.get_tuning()
    try:
        runs = [dict(batch_graph=1), dict(batch_graph=1),        # capture, then replay
                dict(batch_graph=2, batch_tile_w=32, batch_chunks=1),
                # several sample indices per launch (graph and direct launches)
                dict(batch_graph=1, batch_slots=2), dict(batch_graph=2, batch_slots=3),
                dict(batch_graph=1, batch_slots=64, batch_grid=5, batch_waves=8),
                dict(batch_graph=1, batch_tile_w=8, batch_grid=7, batch_chunks=8, batch_stack=2,
                     batch_waves=6),
                dict(batch_graph=1, batch_tile_w=64, batch_grid=4096, batch_chunks=3,
                     batch_waves=5),
                dict(batch_graph=2, batch_waves=1, batch_slots=1)]
        for kw in runs:
            ctx.set_tuning(**kw)
            for s in sensors:
                s.clear()
            st = batch.render(cams, spp, CFG, first_sample=first)
            assert st["rays"] > 0
            assert _same(_films(sensors), solo), kw
    finally:
        ctx.set_tuning(**saved)


def test_batch_graph_replays_with_eager_work_between(gpu):
    items = _small_batch(n=4, seed=9)
    scene_list = [cr.Scene(d) for d, _ in items]
    camds = [c for _, c in items]
    spp = [2, 1, 3, 2]
    solo_a = _solo(scene_list, camds, spp, 5)
    solo_b = _solo(scene_list, camds, spp, 6)
    batch = Batch(scene_list)
    sensors = [cr.Sensor(c["size"]) for c in camds]
    cams = [scenes.make_camera(c, s) for c, s in zip(camds, sensors)]
    ctx = cr.get_context(0)
    saved = ctx.get_tuning()
    try:
        ctx.set_tuning(batch_graph=1)
        for rep in range(3):
            for s in sensors:
                s.clear()
            batch.render(cams, spp, CFG, first_sample=5)
            assert _same(_films(sensors), solo_a), rep
            # eager work on the same stream between replays
            cr.PathTraceIntegrator(CFG).dispatch(scene_list[0], 1, first_sample=999)
            for s in sensors:
                s.clear()
            batch.render(cams, spp, CFG, first_sample=6)   # new first sample: recapture
            assert _same(_films(sensors), solo_b), rep
    finally:
        ctx.set_tuning(**saved)


def test_batch_autotuner_keeps_images(gpu):
    items = _small_batch(n=3, seed=11)
    scene_list = [cr.Scene(d) for d, _ in items]
    camds = [c for _, c in items]
    batch = Batch(scene_list)
    sensors = [cr.Sensor(c["size"]) for c in camds]
    cams = [scenes.make_camera(c, s) for c, s in zip(camds, sensors)]
    spp = [c["spp"] for c in camds]
    ctx = cr.get_context(0)
    saved = ctx.get_tuning()
    try:
        res = autotune.tune_batch(batch, cams, spp, CFG, tile_ws=(8, 32), grids=(64, 1024),
                                  chunks=(1, 8), stacks=(1, 2), reps=1, slots=(1, 2, 8),
                                  slot_spp=[k + 3 for k in spp])
        assert len(res["timings_ms"]) == 16 and len(res["slots_ms"]) == 3
        t = ctx.get_tuning()
        assert t["batch_tile_w"] == res["batch_tile_w"] and t["batch_grid"] == res["batch_grid"]
        assert t["batch_slots"] == res["batch_slots"]
    finally:
        ctx.set_tuning(**saved)


def test_batch_argument_errors(gpu):
    items = _small_batch(n=2, seed=13)
    scene_list = [cr.Scene(d) for d, _ in items]
    batch = Batch(scene_list)
    camds = [c for _, c in items]
    s0 = cr.Sensor(camds[0]["size"])
    cams = [scenes.make_camera(camds[0], s0), scenes.make_camera(camds[1], s0)]
    with pytest.raises(cr.RuntimeError):          # the same film twice
        batch.render(cams, [1, 1], CFG)
    s1 = cr.Sensor(camds[1]["size"])
    cams[1] = scenes.make_camera(camds[1], s1)
    with pytest.raises(cr.RuntimeError):          # an unknown integrator mode
        batch.render(cams, [1, 1], cr.Configs(mode=7))
    with pytest.raises(cr.RuntimeError):
        cr.get_context(0).set_tuning(batch_tile_w=12)


def test_batch_spectral_equals_solo_renders(gpu):
    """Spectral transport in the batched launch (each scene's spectral table
    is part of the batch pool): every XYZ film is bit-identical to the scene
    rendered alone with the spectral megakernel, with one and several sample
    indices per launch, directly and from the captured graph."""
    items = _small_batch(n=4, seed=9)
    scene_list = [cr.Scene(d) for d, _ in items]
    camds = [c for _, c in items]
    spp = [c["spp"] for c in camds]
    cfg = cr.Configs(maxDepth=12, rrThreshold=0.05, useMIS=True, rrStartDepth=3, spectral=True)
    first = 31
    solo = []
    for sc, camd, k in zip(scene_list, camds, spp):
        sensor = cr.Sensor(camd["size"])
        sc.setCamera(scenes.make_camera(camd, sensor))
        cr.PathTraceIntegrator(cfg).dispatch(sc, k, first_sample=first)
        solo.append((sensor.data.copy(), sensor.weight.copy()))
    # the spectral films differ from the RGB ones (XYZ of three hero lanes)
    rgb = _solo(scene_list, camds, spp, first)
    assert not _same(solo, rgb)
    batch = Batch(scene_list)
    sensors = [cr.Sensor(c["size"]) for c in camds]
    cams = [scenes.make_camera(c, s) for c, s in zip(camds, sensors)]
    ctx = cr.get_context(0)
    saved = ctx.get_tuning()
    try:
        for kw in (dict(batch_graph=1), dict(batch_graph=1), dict(batch_graph=2, batch_slots=2),
                   dict(batch_graph=1, batch_slots=64, batch_stack=2, batch_waves=6)):
            ctx.set_tuning(**kw)
            for s in sensors:
                s.clear()
            batch.render(cams, spp, cfg, first_sample=first)
            assert _same(_films(sensors), solo), kw
    finally:
        ctx.set_tuning(**saved)


@pytest.mark.parametrize("spectral", [False, True])
def test_batch_wavefront_equals_solo_renders(gpu, spectral):
    """The batched wavefront: one launch per stage and bounce covers the
    paths of every active scene (each path keeps its scene's index in its
    flags; views and cameras come from the batch table), several sample
    indices per pass as the path state allows, scenes with fewer samples
    dropping out; each film is bit-identical to the scene rendered alone by
    the megakernel, directly and from the captured graph, for several path
    pools (queue capacities) and finishing depths."""
    items = _small_batch(n=5, seed=13)
    scene_list = [cr.Scene(d) for d, _ in items]
    camds = [c for _, c in items]
    spp = [c["spp"] for c in camds]
    spp[2] = 0                          # a scene with nothing to render this call
    cfg = cr.Configs(maxDepth=12, rrThreshold=0.05, useMIS=True, rrStartDepth=3,
                     spectral=spectral)
    first = 41
    solo = []
    for sc, camd, k in zip(scene_list, camds, spp):
        sensor = cr.Sensor(camd["size"])
        sc.setCamera(scenes.make_camera(camd, sensor))
        if k:
            cr.PathTraceIntegrator(cfg).dispatch(sc, k, first_sample=first)
        else:
            sensor.readbackDeviceData()
        solo.append((sensor.data.copy(), sensor.weight.copy()))
    batch = Batch(scene_list)
    sensors = [cr.Sensor(c["size"]) for c in camds]
    cams = [scenes.make_camera(c, s) for c, s in zip(camds, sensors)]
    npix = sum(c["size"][0] * c["size"][1] for c in camds)
    wcfg = cr.Configs(**{**cfg.__dict__, "mode": L.MODE_WAVEFRONT})
    ctx = cr.get_context(0)
    saved = ctx.get_tuning()
    try:
        for kw in (dict(batch_graph=1, wf_pool=npix, wf_finish=65),
                   dict(batch_graph=1, wf_pool=npix, wf_finish=65),     # graph replay
                   dict(batch_graph=2, wf_pool=3 * npix, wf_finish=2),
                   dict(batch_graph=1, wf_pool=2 * npix, wf_finish=1, wf_stack=1),
                   dict(batch_graph=2, wf_pool=1 << 22, wf_finish=3, wf_stack=3,
                        wf_trace_block=64, wf_trace_grid=7, wf_shade_grid=5)):
            ctx.set_tuning(**kw)
            for s in sensors:
                s.clear()
            st = batch.render(cams, spp, wcfg, first_sample=first)
            assert st["rays"] > 0
            assert _same(_films(sensors), solo), kw
    finally:
        ctx.set_tuning(**saved)
