"""Launch tuning (stack placement, block and grid sizes) never changes
results; the autotuner checks that itself and applies the fastest choice."""
import numpy as np
import pytest

import craystal_amd as cr
from craystal_amd import _lib as L
from craystal_amd import autotune, scenes

pytestmark = pytest.mark.gpu


def test_trace_tuning_invariance_and_autotune(gpu):
    data, cam = scenes.cornell_box()
    scene = cr.Scene(data)
    rays = scenes.mixed_batch(data, cam, 50000)
    ctx = cr.get_context(0)
    saved = ctx.get_tuning()
    try:
        # every candidate (triangles per step included) must give the same hits
        res = autotune.tune_trace(scene, rays, reps=2, persistent=((2, 0), (1, 8), (1, 32)),
                                  pt_grids=(64, 1024), tris=(1, 3))
        assert res["stack"] in (1, 2) and res["trace_block"] in (64, 128, 256)
        assert len(res["timings_ms"]) == 6 * 9
        assert ctx.get_tuning()["trace_block"] == res["trace_block"]
        # any-hit: private stack and persistent refill equal the LDS stack
        ctx.set_tuning(stack=1, persistent=2)
        a = scene.trace_any(rays)
        ctx.set_tuning(stack=2)
        b = scene.trace_any(rays)
        ctx.set_tuning(stack=1, persistent=1, refill=4, pt_grid=32)
        c = scene.trace_any(rays)
        ctx.set_tuning(trace_tris=2)
        d = scene.trace_any(rays)
        assert np.array_equal(a, b) and np.array_equal(a, c) and np.array_equal(a, d)
    finally:
        ctx.set_tuning(**saved)


def test_render_tuning_invariance(gpu):
    data, camd = scenes.cornell_box()
    scene = cr.Scene(data)
    sensor = cr.Sensor((64, 64))
    cam = scenes.make_camera(camd, sensor)
    scene.setCamera(cam)
    cfg = cr.Configs(maxDepth=8, rrThreshold=0.05, useMIS=True, rrStartDepth=3)
    ctx = cr.get_context(0)
    saved = ctx.get_tuning()
    try:
        # every candidate (register budgets included) must give the same film, bit for bit
        res = autotune.tune_render(scene, cfg, spp=1, wf_grids=(64, 512), wf_blocks=(64, 256),
                                   render_bands=(1, 2), render_waves=(1, 4, 5, 6, 8))
        # stacks x budgets x (2 bands + 2 regeneration grids)
        assert len(res["megakernel_ms"]) == 2 * 5 * 4
        assert res["mode"] in (L.MODE_MEGAKERNEL, L.MODE_WAVEFRONT)
        assert len(res["wavefront_ms"]) == 16
        t = ctx.get_tuning()
        assert t["render_stack"] == res["render_stack"] and t["wf_stack"] == res["wf_stack"]
    finally:
        ctx.set_tuning(**saved)


def test_tuning_rejects_bad_values(gpu):
    ctx = cr.get_context(0)
    with pytest.raises(cr.RuntimeError):
        ctx.set_tuning(trace_block=96)
    with pytest.raises(cr.RuntimeError):
        ctx.set_tuning(stack=7)
    with pytest.raises(cr.RuntimeError):
        ctx.set_tuning(wf_refill=65)
    with pytest.raises(cr.RuntimeError):
        ctx.set_tuning(batch_slots=65)
    with pytest.raises(cr.RuntimeError):
        ctx.set_tuning(trace_tris=5)
    for w in (2, 3, 7):
        with pytest.raises(cr.RuntimeError):
            ctx.set_tuning(render_waves=w)
        with pytest.raises(cr.RuntimeError):
            ctx.set_tuning(batch_waves=w)
