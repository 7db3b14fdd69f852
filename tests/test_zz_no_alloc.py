"""crh_trace_closest_device_noalloc, the closest-hit query for callers that
cannot allocate (no reference counterpart), against crh_trace_closest_device
on the same inputs: hits bit-identical and node / triangle totals equal, on
two calls in a row whose second batch is smaller (the shared scratch array
still holds the first call's contents), with the counting kernel (work counts
taken in chunks of 2^20 rays through the scratch array) and with the
persistent kernel (its dispenser counter in the scratch array)."""
import ctypes as C

import numpy as np
import pytest

import craystal_amd as cr
from craystal_amd import _lib as L
from craystal_amd import scenes

pytestmark = pytest.mark.gpu


def _noalloc(scene, rays, n, hits, count):
    st = L.crh_stats()
    L.call("crh_trace_closest_device_noalloc", scene.handle, rays.handle, n, hits.handle,
           1 if count else 0, C.byref(st))
    return st


def test_noalloc_trace_matches_trace_closest_device(gpu):
    data, cam = scenes.spheres_scene(n_spheres=12, subdiv=4)
    gs = cr.Scene(data)
    ctx = cr.get_context(0)
    saved = ctx.get_tuning()
    big = 1_300_000                     # more than one 2^20-ray chunk of work counts
    rays = scenes.mixed_batch(data, cam, big)
    rb = cr.DeviceBuffer(rays.nbytes)
    rb.copyFromHost(rays)
    hb_ref = cr.DeviceBuffer(big * 16)
    hb = cr.DeviceBuffer(big * 16)
    try:
        for persistent in (2, 1):
            ctx.set_tuning(persistent=persistent)
            for count in (True, False):
                for n in (big, 200_003):     # second call smaller
                    print(f"noalloc persistent={persistent} count={count} n={n}", flush=True)
                    ref = cr.trace_closest_device(gs, rb, n, hb_ref, count_work=count)
                    st = _noalloc(gs, rb, n, hb, count)
                    a = np.zeros(n, L.HIT_DTYPE)
                    b = np.zeros(n, L.HIT_DTYPE)
                    hb_ref.copyToHost(a)
                    hb.copyToHost(b)
                    assert a.tobytes() == b.tobytes()
                    assert st.rays == n
                    if count:
                        assert st.node_visits == ref["node_visits"] > 0
                        assert st.tri_tests == ref["tri_tests"] > 0
        assert (b["prim"] != L.CRH_INVALID_PRIM).mean() > 0.2
    finally:
        ctx.set_tuning(**saved)


def test_noalloc_trace_argument_errors(gpu):
    data, cam = scenes.cornell_box()
    gs = cr.Scene(data)
    rb = cr.DeviceBuffer(32 * 32)
    hb = cr.DeviceBuffer(16 * 16)
    with pytest.raises(L.RuntimeError, match="hit buffer too small"):
        _noalloc(gs, rb, 32, hb, True)
    with pytest.raises(L.RuntimeError, match="ray buffer too small"):
        _noalloc(gs, rb, 64, cr.DeviceBuffer(64 * 16), False)
