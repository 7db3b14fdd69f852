"""The GPU binned-SAH builder (crh_bvh_build_device) against the host builder,
which is itself bit-exact with the oracle's restatement of BVH::build
(Core/BVH.cu:88-338): node arrays (tags, children, boxes to the bit),
primitive order and statistics identical, on the committed fixtures, random
and degenerate soups (coincident centroids, zero-area triangles, NaN
vertices, mixed +0 / -0 coordinates), several parameter sets, the
983k-triangle scene (its SHA-256 fixture) and the 43.8M-triangle terrain."""
import hashlib
import json
import time
from pathlib import Path

import numpy as np
import pytest

import craystal_amd as cr
from craystal_amd import scenes

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"
STATS = ("node_count", "max_depth", "max_leaf", "root", "prim_count", "internal_count",
         "root_min", "root_max")


def _pair(P, I, params=(32, 4, 20)):
    h = cr.BVH(P, I, *params)
    d = cr.BVH(P, I, *params, device=True)
    return h, d


def _assert_same(h, d):
    nh, ph = h.export()
    nd, pd = d.export()
    assert nh.tobytes() == nd.tobytes(), "node arrays differ"
    assert np.array_equal(ph, pd), "primitive orders differ"
    sh, sd = h.stats(), d.stats()
    for k in STATS:
        assert sh[k] == sd[k], (k, sh[k], sd[k])


def _ico():
    mat = scenes.diffuse((0.5, 0.5, 0.5))
    return cr.SceneData(meshes=[scenes.icosphere((0, 0, 0), 1.0, 3, 0),
                                scenes.icosphere((1.5, 0.2, 0.1), 0.7, 2, 0)],
                        materials=[mat])


def test_device_build_equals_golden_fixtures(gpu):
    g = np.load(GOLD / "bvh_cornell.npz")
    n, p = cr.BVH(*scenes.cornell_box()[0].concatenated(), device=True).export()
    assert n.tobytes() == g["nodes"].tobytes() and np.array_equal(p, g["prims"])
    g = np.load(GOLD / "bvh_ico1k.npz")
    n, p = cr.BVH(*_ico().concatenated(), device=True).export()
    assert n.tobytes() == g["nodes"].tobytes() and np.array_equal(p, g["prims"])


@pytest.mark.parametrize("params", [(32, 4, 20), (8, 1, 64), (64, 2, 12), (2, 4, 20),
                                    (16, 8, 3), (32, 1, 40)])
def test_device_build_equals_host_on_params(gpu, params):
    for data in (scenes.cornell_box()[0], _ico(), scenes.batch_scene(3)[0]):
        _assert_same(*_pair(*data.concatenated(), params))


def test_device_build_degenerate_and_random_soups(gpu):
    rng = np.random.default_rng(11)
    z = np.float32(0.0)
    signed = rng.normal(size=(3000, 3)).astype(np.float32)
    signed[rng.random((3000, 3)) < 0.3] = z
    signed[rng.random((3000, 3)) < 0.3] = -z          # mixed +0 / -0 coordinates
    nanp = rng.normal(size=(500, 3)).astype(np.float32)
    nanp[rng.integers(0, 500, 20), rng.integers(0, 3, 20)] = np.nan
    cases = [
        (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]] * 10, np.float32), np.arange(30)),
        (rng.normal(size=(300, 3)).astype(np.float32), rng.integers(0, 300, 600)),
        (np.round(rng.uniform(-1, 1, (400, 3)), 1).astype(np.float32),
         rng.integers(0, 400, 30000)),                 # coincident centroids, large tasks
        (np.zeros((3, 3), np.float32), np.array([0, 1, 2])),
        (signed, rng.integers(0, 3000, 60000)),
        (nanp, rng.integers(0, 500, 9000)),
        (np.zeros((0, 3), np.float32), np.zeros(0, np.uint32)),
        (rng.normal(size=(50000, 3)).astype(np.float32) * np.float32(1e-30),
         rng.integers(0, 50000, 150000)),              # tiny extents
    ]
    for P, I in cases:
        _assert_same(*_pair(P, I.astype(np.uint32)))
        _assert_same(*_pair(P, I.astype(np.uint32), (8, 2, 64)))


def test_device_build_spheres_1m_hash(gpu):
    meta = json.loads((GOLD / "bvh_spheres1m.json").read_text())
    P, I = scenes.spheres_scene()[0].concatenated()
    d = cr.BVH(P, I, device=True)
    n, p = d.export()
    h = hashlib.sha256()
    h.update(n.tobytes())
    h.update(p.tobytes())
    assert h.hexdigest() == meta["sha256"]
    st = d.stats()
    assert st["node_count"] == meta["node_count"] and st["max_leaf"] == meta["max_leaf"]


def test_device_build_scene_renders_identically(gpu):
    """A scene whose BVH is built on the GPU traces and renders exactly like
    the host-built one."""
    data, camd = scenes.cornell_box()
    rays = scenes.mixed_batch(data, camd, 50000)
    a = cr.Scene(data)
    data.bvhDevice = True
    b = cr.Scene(data)
    assert a.trace_closest(rays).tobytes() == b.trace_closest(rays).tobytes()


def test_device_build_terrain_43m(gpu):
    data, _ = scenes.terrain_scene()
    P, I = data.concatenated()
    t0 = time.perf_counter()
    d = cr.BVH(P, I, device=True)
    td = time.perf_counter() - t0
    t0 = time.perf_counter()
    h = cr.BVH(P, I)
    th = time.perf_counter() - t0
    print(f"terrain {len(I) // 3} triangles: GPU build {td:.3f} s "
          f"({d.stats()['build_ms']:.1f} ms in the call), host build {th:.3f} s", flush=True)
    _assert_same(h, d)
