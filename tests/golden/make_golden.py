"""Writes the committed golden fixtures under tests/golden/ from the CPU
oracle (oracle/oracle.cpp) and the seeded scene generators.

Fixtures freeze oracle outputs that the independent pins in
tests/test_oracle_pins.py have checked, so later changes to either side of
the boundary are compared with fixed data:
  sampler.json         raw Sampler outputs for a few (x, y, sample) keys
  bvh_cornell.npz      reference-layout node array + primitive order, 36 tris
  bvh_ico1k.npz        same for a 1,280-triangle two-icosphere soup
  bvh_spheres1m.json   SHA-256 of node array + primitive order, 983,044 tris
  hits_cornell.npz     4,096-ray mixed batch + closest hits
  hits_spheres1m.json  SHA-256 of the closest hits of the 16,777,216-ray batch of the
                       983,044-triangle scene (8M primary + 8M secondary, ray seed
                       0x5EED), of the whole records and of (prim, t) alone
  hits_spheres1m_sample.npz  every 256th of those hits (65,536 records)
Run: python tests/golden/make_golden.py [spheres-hits]
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

import craystal_amd as cr  # noqa: E402
import oracle  # noqa: E402
from craystal_amd import scenes  # noqa: E402


def ico1k():
    mat = scenes.diffuse((0.5, 0.5, 0.5))
    return cr.SceneData(meshes=[scenes.icosphere((0, 0, 0), 1.0, 3, 0),
                                scenes.icosphere((1.5, 0.2, 0.1), 0.7, 2, 0)],
                        materials=[mat])


def digest(nodes, prims):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(nodes).tobytes())
    h.update(np.ascontiguousarray(prims).tobytes())
    return h.hexdigest()


SPHERES_RAYS = 1 << 24
SAMPLE_STRIDE = 256


def prim_t_bytes(hits):
    """(prim, t bits) of every hit as consecutive uint32 pairs."""
    return np.column_stack((hits["prim"], hits["t"].view(np.uint32))).astype(np.uint32).tobytes()


def spheres_hits():
    d, cam = scenes.spheres_scene()
    rays = scenes.mixed_batch(d, cam, SPHERES_RAYS)
    hits = oracle.Scene(d).trace_closest(rays)
    (HERE / "hits_spheres1m.json").write_text(json.dumps(
        {"rays": SPHERES_RAYS, "ray_seed": scenes.RAY_SEED,
         "sha256_records": hashlib.sha256(hits.tobytes()).hexdigest(),
         "sha256_prim_t": hashlib.sha256(prim_t_bytes(hits)).hexdigest(),
         "hit_fraction": float((hits["prim"] != 0xFFFFFFFF).mean()),
         "sample_stride": SAMPLE_STRIDE}, indent=1))
    np.savez_compressed(HERE / "hits_spheres1m_sample.npz", hits=hits[::SAMPLE_STRIDE])


def main():
    if sys.argv[1:] == ["spheres-hits"]:
        spheres_hits()
        return
    keys = [(0, 0, 0), (123, 456, 0), (123, 456, 7), (1919, 1079, 63)]
    seqs = {f"{x},{y},{i}": oracle.sampler_seq(x, y, i, 16)[1].tolist() for x, y, i in keys}
    (HERE / "sampler.json").write_text(json.dumps(seqs, indent=1))

    d, cam = scenes.cornell_box()
    P, I = d.concatenated()
    nodes, prims, _ = oracle.build_bvh(P, I)
    np.savez(HERE / "bvh_cornell.npz", nodes=nodes, prims=prims)
    rays = scenes.mixed_batch(d, cam, 4096)
    hits = oracle.Scene(d).trace_closest(rays)
    np.savez(HERE / "hits_cornell.npz", rays=rays, hits=hits)

    d2 = ico1k()
    P2, I2 = d2.concatenated()
    n2, p2, _ = oracle.build_bvh(P2, I2)
    np.savez(HERE / "bvh_ico1k.npz", nodes=n2, prims=p2)

    d3, _ = scenes.spheres_scene()
    P3, I3 = d3.concatenated()
    n3, p3, st = oracle.build_bvh(P3, I3)
    (HERE / "bvh_spheres1m.json").write_text(json.dumps(
        {"triangles": int(len(I3) // 3), "node_count": int(len(n3)), "max_depth": st["max_depth"],
         "max_leaf": st["max_leaf"], "sha256": digest(n3, p3)}, indent=1))
    spheres_hits()


if __name__ == "__main__":
    main()
