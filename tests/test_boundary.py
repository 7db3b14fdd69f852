"""The C-ABI boundary on the CPU: the library loads and exports every symbol
include/crayhip.h declares; host-side entry points (BVH build, material
finalize, camera update, argument validation) behave like the reference's
host code; nothing here launches a kernel."""
import hashlib
import json
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import craystal_amd as cr
from craystal_amd import _lib as L
from craystal_amd import scenes

ROOT = Path(__file__).resolve().parents[1]
GOLD = ROOT / "tests" / "golden"


def _declared():
    text = (ROOT / "include" / "crayhip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(crh_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = cr.load()
    names = _declared()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", str(L.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (crh_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(lib, n)
    assert set(L.SIGNATURES) | {"crh_last_error"} >= set(names)


def test_library_is_gfx950_and_no_cpu_fallback_symbols():
    out = subprocess.run(["strings", str(L.LIB_PATH)], capture_output=True, text=True).stdout
    assert "gfx950" in out
    assert "orc_" not in out  # the product never contains the oracle


def _bvh_pair(P, I, params=(32, 4, 20)):
    import oracle
    b = cr.BVH(P, I, *params)
    n1, p1 = b.export()
    n2, p2, st = oracle.build_bvh(P, I, *params)
    return n1, p1, n2, p2, b.stats(), st


@pytest.mark.parametrize("params", [(32, 4, 20), (8, 1, 64), (64, 2, 12), (2, 4, 20)])
def test_product_bvh_equals_oracle(orc, params):
    for data in (scenes.cornell_box()[0], _ico()):
        P, I = data.concatenated()
        n1, p1, n2, p2, s1, s2 = _bvh_pair(P, I, params)
        assert n1.tobytes() == n2.tobytes()
        assert np.array_equal(p1, p2)
        for k in ("node_count", "max_depth", "max_leaf", "root"):
            assert s1[k] == s2[k]


def _ico():
    mat = scenes.diffuse((0.5, 0.5, 0.5))
    return cr.SceneData(meshes=[scenes.icosphere((0, 0, 0), 1.0, 3, 0),
                                scenes.icosphere((1.5, 0.2, 0.1), 0.7, 2, 0)],
                        materials=[mat])


def test_product_bvh_equals_golden_fixtures():
    g = np.load(GOLD / "bvh_cornell.npz")
    P, I = scenes.cornell_box()[0].concatenated()
    n, p = cr.BVH(P, I).export()
    assert n.tobytes() == g["nodes"].tobytes() and np.array_equal(p, g["prims"])
    g = np.load(GOLD / "bvh_ico1k.npz")
    P, I = _ico().concatenated()
    n, p = cr.BVH(P, I).export()
    assert n.tobytes() == g["nodes"].tobytes() and np.array_equal(p, g["prims"])


def test_product_bvh_spheres_1m_hash():
    meta = json.loads((GOLD / "bvh_spheres1m.json").read_text())
    P, I = scenes.spheres_scene()[0].concatenated()
    b = cr.BVH(P, I)
    n, p = b.export()
    h = hashlib.sha256()
    h.update(n.tobytes())
    h.update(p.tobytes())
    assert h.hexdigest() == meta["sha256"]
    st = b.stats()
    assert st["node_count"] == meta["node_count"] and st["max_leaf"] == meta["max_leaf"]


def test_product_bvh_degenerate_and_random_soups(orc):
    rng = np.random.default_rng(11)
    cases = [
        (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]] * 10, np.float32), np.arange(30)),
        (rng.normal(size=(300, 3)).astype(np.float32), rng.integers(0, 300, 600)),
        (np.round(rng.uniform(-1, 1, (400, 3)), 1).astype(np.float32),
         rng.integers(0, 400, 3000)),              # many coincident centroids
        (np.zeros((3, 3), np.float32), np.array([0, 1, 2])),   # zero-area triangle
    ]
    for P, I in cases:
        I = I.astype(np.uint32)
        n1, p1, n2, p2, *_ = _bvh_pair(P, I)
        assert n1.tobytes() == n2.tobytes() and np.array_equal(p1, p2)


def test_material_finalize_semantics():
    m = cr.MaterialData(emission=(0, 0, 1.0)).finalize()          # Core/Material.h:57-65
    assert m.flags == (L.MATF_EMISSIVE | L.MATF_DIFFUSE)
    m = cr.MaterialData(specularRefl=(0, 0.1, 0)).finalize()
    assert m.flags == 0


def test_argument_errors_raise_runtime_error():
    with pytest.raises(cr.RuntimeError) as e:
        cr.BVH(np.zeros((3, 3), np.float32), np.array([0, 1, 5], np.uint32))
    assert e.value.code == L.CRH_ERR_INVALID_ARG and "out of range" in str(e.value)
    with pytest.raises(cr.RuntimeError):
        cr.BVH(np.zeros((3, 3), np.float32), np.array([0, 1, 2], np.uint32), bins=1)
    with pytest.raises(cr.RuntimeError):
        L.call("crh_material_finalize", None)
    n = np.zeros(1, np.uint32)
    assert L.load().crh_version(n.ctypes.data_as(L.C.POINTER(L.u32))) == 0 and n[0] >= 1


def test_context_on_missing_device_raises():
    """On a host without a GPU (or with an out-of-range index) the context
    call fails loudly with a HIP-shaped message instead of falling back."""
    with pytest.raises(cr.RuntimeError) as e:
        cr.core.Context(device=4096)
    assert e.value.code in (L.CRH_ERR_INVALID_ARG, L.CRH_ERR_HIP)


def test_scene_generators_are_deterministic():
    a, _ = scenes.spheres_scene(n_spheres=6, subdiv=1)
    b, _ = scenes.spheres_scene(n_spheres=6, subdiv=1)
    Pa, Ia = a.concatenated()
    Pb, Ib = b.concatenated()
    assert Pa.tobytes() == Pb.tobytes() and Ia.tobytes() == Ib.tobytes()
    d, cam = scenes.cornell_box()
    assert d.triangle_count == 36 and d.emissiveIndex == [34, 35]
    r1 = scenes.mixed_batch(d, cam, 1000)
    r2 = scenes.mixed_batch(d, cam, 1000)
    assert r1.tobytes() == r2.tobytes()
    assert np.allclose(np.linalg.norm(r1["dir"], axis=1), 1, atol=1e-5)


def test_oracle_reproduces_committed_hit_fixtures(orc):
    """The oracle's closest hits against the committed fixtures: the Cornell
    batch whole, and every 256th ray of the 16,777,216-ray batch of the
    983k-triangle scene (its full SHA-256 is checked on the GPU)."""
    g = np.load(GOLD / "hits_cornell.npz")
    data, _ = scenes.cornell_box()
    assert orc.Scene(data).trace_closest(g["rays"]).tobytes() == g["hits"].tobytes()
    meta = json.loads((GOLD / "hits_spheres1m.json").read_text())
    data, cam = scenes.spheres_scene()
    rays = scenes.mixed_batch(data, cam, meta["rays"])[::meta["sample_stride"]]
    sample = np.load(GOLD / "hits_spheres1m_sample.npz")["hits"]
    assert orc.Scene(data).trace_closest(rays).tobytes() == sample.tobytes()
    assert 0.3 < (sample["prim"] != 0xFFFFFFFF).mean() < 0.4
