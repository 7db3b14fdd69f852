"""GPU parity of the ray queries against the oracle (bit-exact hit IDs and t).

Reference semantics: BVHData::intersect (Core/BVH.cu:13-84), intersectShape
(Scene/Shape.h:17-37), Moller-Trumbore (Scene/Triangle.cu:40-75) and
intersectOcclusion (Scene/Scene.cu:16-29).
"""
import numpy as np
import pytest

import craystal_amd as cr
from craystal_amd import _lib as L
from craystal_amd import scenes

pytestmark = pytest.mark.gpu


def _assert_hits_equal(h_gpu, h_orc):
    assert np.array_equal(h_gpu["prim"], h_orc["prim"]), \
        f"prim mismatch at {np.nonzero(h_gpu['prim'] != h_orc['prim'])[0][:10]}"
    # t, u, v bit-exact (the path uses only + - * / and compares)
    for k in ("t", "u", "v"):
        a = h_gpu[k].view(np.uint32)
        b = h_orc[k].view(np.uint32)
        assert np.array_equal(a, b), f"{k} mismatch at {np.nonzero(a != b)[0][:10]}"


def _scene_pair(data, orc):
    return cr.Scene(data), orc.Scene(data)


def test_cornell_closest_and_any(gpu, orc):
    data, cam = scenes.cornell_box()
    gs, os_ = _scene_pair(data, orc)
    rays = scenes.mixed_batch(data, cam, 20000)
    _assert_hits_equal(gs.trace_closest(rays), os_.trace_closest(rays))
    short = rays.copy()
    short["tmax"] = np.float32(0.7)
    assert np.array_equal(gs.trace_any(short), os_.trace_any(short))
    assert np.array_equal(gs.trace_any(rays), os_.trace_any(rays))


def test_icosphere_1k_and_bvh_params(gpu, orc):
    mat = scenes.diffuse((0.5, 0.5, 0.5))
    for params in [(32, 4, 20), (8, 1, 64), (16, 2, 6)]:
        data = cr.SceneData(meshes=[scenes.icosphere((0, 0, 0), 1.0, 3, 0),
                                    scenes.icosphere((1.5, 0.2, 0.1), 0.7, 2, 0)],
                            materials=[mat], bvhParams=params)
        gs, os_ = _scene_pair(data, orc)
        assert gs.bvh_stats()["node_count"] == os_.bvh_stats()["node_count"]
        cam = dict(pos=(0, 0.5, 4), target=(0, 0, 0), up=(0, 1, 0), fov=50, size=(64, 64))
        rays = scenes.mixed_batch(data, cam, 8192)
        _assert_hits_equal(gs.trace_closest(rays), os_.trace_closest(rays))


def test_axis_parallel_and_degenerate_rays(gpu, orc):
    """Flat boxes, axis-parallel directions (1/d = inf, 0*inf = NaN in the
    slab test), rays starting on planes, zero-length and tiny tmax."""
    data, cam = scenes.cornell_box()
    gs, os_ = _scene_pair(data, orc)
    rng = np.random.default_rng(7)
    n = 4096
    rays = np.zeros(n, L.RAY_DTYPE)
    o = rng.uniform(-1.2, 2.2, (n, 3)).astype(np.float32)
    o[: n // 4, 1] = 0.0          # on the floor plane
    o[n // 4: n // 2, 0] = -1.0   # on the left wall plane
    d = np.zeros((n, 3), np.float32)
    axis = rng.integers(0, 3, n)
    sign = np.where(rng.uniform(size=n) < 0.5, -1.0, 1.0).astype(np.float32)
    d[np.arange(n), axis] = sign
    d[n // 2:] = rng.normal(size=(n - n // 2, 3)).astype(np.float32)
    d[-64:] = 0.0                 # zero direction
    rays["origin"] = o
    rays["dir"] = d
    rays["tmin"] = 0.0
    rays["tmax"] = np.where(rng.uniform(size=n) < 0.2, np.float32(0.5), np.float32(np.inf))
    _assert_hits_equal(gs.trace_closest(rays), os_.trace_closest(rays))
    assert np.array_equal(gs.trace_any(rays), os_.trace_any(rays))


def test_empty_scene_and_empty_batch(gpu, orc):
    data = cr.SceneData(meshes=[], materials=[scenes.diffuse((1, 1, 1))])
    gs = cr.Scene(data)
    rays = scenes.primary_rays(dict(pos=(0, 0, 3), target=(0, 0, 0), up=(0, 1, 0), fov=40,
                                    size=(8, 8)), 100)
    h = gs.trace_closest(rays)
    assert (h["prim"] == L.CRH_INVALID_PRIM).all() and np.isinf(h["t"]).all()
    assert not gs.trace_any(rays).any()
    assert len(gs.trace_closest(rays[:0])) == 0


def test_spheres_1m_batch_device_resident(gpu, orc):
    """983k-triangle scene: a 1M-ray mixed batch through the device-resident
    entry point, hit IDs and t bit-exact, node/triangle work counts equal to
    the oracle's reference-order traversal."""
    data, cam = scenes.spheres_scene()
    gs, os_ = _scene_pair(data, orc)
    n = 1 << 20
    rays = scenes.mixed_batch(data, cam, n)
    rb = cr.DeviceBuffer(rays.nbytes)
    hb = cr.DeviceBuffer(n * 16)
    rb.copyFromHost(rays)
    st = cr.trace_closest_device(gs, rb, n, hb, count_work=True)
    hits = np.zeros(n, L.HIT_DTYPE)
    hb.copyToHost(hits)
    oh, nv, nt = os_.trace_closest(rays, counts=True)
    _assert_hits_equal(hits, oh)
    assert st["node_visits"] == int(nv.sum(dtype=np.uint64))
    assert st["tri_tests"] == int(nt.sum(dtype=np.uint64))
    assert (hits["prim"] != L.CRH_INVALID_PRIM).mean() > 0.3


def test_persistent_refill_bit_identical(gpu, orc):
    """Lanes refilling with new rays (persistent kernels) give the same hits
    as one ray per thread, on the 983k-triangle scene and ragged sizes."""
    data, cam = scenes.spheres_scene()
    gs, os_ = _scene_pair(data, orc)
    ctx = cr.get_context(0)
    saved = ctx.get_tuning()
    try:
        for n, refill, grid in [(1 << 18, 16, 2048), (77777, 1, 7), (4099, 64, 3)]:
            rays = scenes.mixed_batch(data, cam, n)
            ctx.set_tuning(persistent=2)
            base = gs.trace_closest(rays)
            ctx.set_tuning(persistent=1, refill=refill, pt_grid=grid)
            pt = gs.trace_closest(rays)
            assert pt.tobytes() == base.tobytes()
            assert np.array_equal(gs.trace_any(rays), os_.trace_any(rays))
        _assert_hits_equal(pt, os_.trace_closest(rays))
    finally:
        ctx.set_tuning(**saved)


def test_slab_fast_path_nan_and_plane_cases(gpu, orc):
    """Boxes without a flat axis take the select-free slab test; it must give
    the reference's hits and node/triangle counts when a slab value is NaN
    (zero direction component with the origin on a box plane: 0 * inf), for
    tMin = -0, and when tMin or tMax is NaN (those rays take the reference
    path for every box)."""
    mat = scenes.diffuse((0.5, 0.5, 0.5))
    data = cr.SceneData(meshes=[scenes.icosphere((0, 0, 0), 1.0, 3, 0),
                                scenes.icosphere((1.7, 0.3, -0.4), 0.6, 2, 0)], materials=[mat])
    gs, os_ = _scene_pair(data, orc)
    P = np.concatenate([m.position for m in data.meshes])
    rng = np.random.default_rng(23)
    n = 1 << 14
    rays = np.zeros(n, L.RAY_DTYPE)
    o = rng.uniform(-2.5, 3.0, (n, 3)).astype(np.float32)
    pick = P[rng.integers(0, len(P), n)]
    on_plane = rng.integers(0, 3, n)
    o[np.arange(n), on_plane] = pick[np.arange(n), on_plane]     # origin on a vertex plane
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d[np.arange(n), on_plane] = 0.0                                # parallel to that plane
    both = rng.uniform(size=n) < 0.3
    d[both, (on_plane[both] + 1) % 3] = 0.0
    rays["origin"] = o
    rays["dir"] = d
    kind = rng.integers(0, 6, n)
    rays["tmin"] = np.select([kind == 1, kind == 2], [np.float32(-0.0), np.float32(np.nan)],
                             np.float32(0.0))
    rays["tmax"] = np.select([kind == 3, kind == 4], [np.float32(np.nan), np.float32(1.5)],
                             np.float32(np.inf))
    oh, nv, nt = os_.trace_closest(rays, counts=True)
    _assert_hits_equal(gs.trace_closest(rays), oh)
    assert np.array_equal(gs.trace_any(rays), os_.trace_any(rays))
    rb, hb = cr.DeviceBuffer(rays.nbytes), cr.DeviceBuffer(n * 16)
    rb.copyFromHost(rays)
    cnt = cr.trace_closest_device(gs, rb, n, hb, count_work=True)
    assert cnt["node_visits"] == int(nv.sum(dtype=np.uint64))
    assert cnt["tri_tests"] == int(nt.sum(dtype=np.uint64))
    assert (oh["prim"] != L.CRH_INVALID_PRIM).sum() > 1000


GOLD = __import__("pathlib").Path(__file__).resolve().parent / "golden"


def _prim_t_bytes(hits):
    return np.column_stack((hits["prim"], hits["t"].view(np.uint32))).astype(np.uint32).tobytes()


def test_committed_hit_fixtures(gpu):
    """The GPU against the oracle's committed outputs: the 4,096-ray Cornell
    batch, and the full 16,777,216-ray batch of the 983k-triangle scene by its
    SHA-256 (whole records and (prim, t)) plus every 256th record."""
    import hashlib
    import json
    g = np.load(GOLD / "hits_cornell.npz")
    data, _ = scenes.cornell_box()
    assert cr.Scene(data).trace_closest(g["rays"]).tobytes() == g["hits"].tobytes()
    meta = json.loads((GOLD / "hits_spheres1m.json").read_text())
    data, cam = scenes.spheres_scene()
    n = meta["rays"]
    rays = scenes.mixed_batch(data, cam, n)
    gs = cr.Scene(data)
    rb, hb = cr.DeviceBuffer(rays.nbytes), cr.DeviceBuffer(n * 16)
    rb.copyFromHost(rays)
    cr.trace_closest_device(gs, rb, n, hb)
    hits = np.zeros(n, L.HIT_DTYPE)
    hb.copyToHost(hits)
    sample = np.load(GOLD / "hits_spheres1m_sample.npz")["hits"]
    assert hits[::meta["sample_stride"]].tobytes() == sample.tobytes()
    assert hashlib.sha256(_prim_t_bytes(hits)).hexdigest() == meta["sha256_prim_t"]
    assert hashlib.sha256(hits.tobytes()).hexdigest() == meta["sha256_records"]
