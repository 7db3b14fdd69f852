#!/usr/bin/env python3
"""Benchmark of the MI355X rendering hot path (one GPU).

Workload: the ~1.0M-triangle scene (48 subdivision-5 icospheres + ground quad
+ 2 x 2 m area light + constant environment, Principled materials),
1920 x 1080, max depth 16, Russian roulette from depth 3, MIS on; rendered by
the plain megakernel.  One step = one sample per pixel over the whole image
(2,073,600 paths).  `value` = Msamples/s over the K timed steps, with the
scene and film already resident in HBM.

Also reported: Grays/s of the render (extension + shadow + MIS rays counted by
the kernel), the closest-hit throughput on a 16,777,216-ray seeded batch
(8M primary + 8M secondary, device resident), the roofline of the dominant
kernel from its HIP-event-timed launches, the wavefront integrator, the
64-scene batch, the guided-vs-plain comparison, the 43.8M-triangle terrain
(four 3840 x 2160 views in one batched launch), and the CPU oracle timed on
the host cores on a bounded sample (spp passes of the same image).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
B_RAY_FIXED = 32 + 16      # ray in + hit out
B_NODE = 64                # one internal visit: both child boxes + child words
B_TRI = 52                 # 48 B index + vertices, 4 B primitive map
B_FILM = 16 * 9            # (3 + 1) floats x up to 9 pixels


def progress(msg: str):
    """One line per finished leg on stderr (the JSON result goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--trace-rays", type=int, default=1 << 24)
    ap.add_argument("--trace-reps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-trace", action="store_true", help="skip the closest-hit batch")
    ap.add_argument("--guided-seconds", type=float, default=6.0,
                    help="render time of each run of the guided-vs-plain comparison (0 = skip)")
    ap.add_argument("--guided-materials", type=int, choices=(1, 2, 3), default=1,
                    help="guided leg: BSDFs guided (1 Lambertian, 2 Principled, 3 both)")
    ap.add_argument("--guided-train-steps", type=int, default=32,
                    help="guided leg: Adam steps per training pass")
    ap.add_argument("--guided-train-batch", type=int, default=16384,
                    help="guided leg: training samples per Adam step")
    ap.add_argument("--guided-runs", type=int, default=9,
                    help="independent runs per method in the guided-vs-plain comparison")
    ap.add_argument("--guided-size", type=int, choices=(256, 1024), default=1024,
                    help="image size of the guided-vs-plain comparison (its 65,536-spp "
                         "reference is committed for both)")
    ap.add_argument("--batch-scenes", type=int, default=64,
                    help="scenes of the many-scene batch leg (0 = skip)")
    ap.add_argument("--no-terrain", action="store_true",
                    help="skip the large-scale leg (the 43.8M-triangle terrain, four "
                         "3840x2160 views in one batched launch, and its closest-hit batch)")
    ap.add_argument("--tuning", type=str, default=None,
                    help="JSON of launch shapes (crh_tuning fields) to apply instead of "
                         "running the autotuner on the main workload")
    ap.add_argument("--only-main", action="store_true",
                    help="only the main workload (no other legs, CPU baseline or traffic runs)")
    ap.add_argument("--no-c1", action="store_true",
                    help="skip the small spectral Cornell-box configuration")
    ap.add_argument("--terrain-bvh", choices=("fast", "reference"), default="fast",
                    help="terrain BVH: built on the GPU without the depth cap (fast) or "
                         "with the reference's parameters")
    ap.add_argument("--terrain-guided-spp", type=int, default=2,
                    help="spp of every terrain view trained, then timed, with the 128-wide "
                         "guide (0 = skip)")
    ap.add_argument("--terrain-spp", type=int, default=4,
                    help="timed spp per view of the terrain leg")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc runs that measure HBM traffic")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="counter-collection mode: one 256 MiB device copy (traffic "
                         "calibration), one 16,777,216-ray closest-hit launch and one render "
                         "pass; prints the launch sizes and exits")
    ap.add_argument("--profile-tuning", default=None,
                    help="with --profile-kernels: JSON of launch-shape fields to apply first")
    ap.add_argument("--profile-wavefront", action="store_true",
                    help="counter-collection mode for the wavefront: the spheres scene at "
                         "1920x1080, 8 spp (with --profile-tuning applied first)")
    ap.add_argument("--profile-terrain", action="store_true",
                    help="counter-collection mode for the terrain: one spp of its four "
                         "3840x2160 views (batched, no graph) and a 16,777,216-ray closest-hit "
                         "batch (with --profile-tuning applied first)")
    ap.add_argument("--profile-guided", action="store_true",
                    help="counter-collection mode for the guiding kernels: the occluded room "
                         "at 256x256, 16 spp of the fused guided kernel with training, 16 "
                         "without, 16 of the guided wavefront, "
                         "then 16 spp of the plain megakernel and of the plain wavefront")
    a = ap.parse_args()
    if a.only_main:
        a.no_cpu = a.no_trace = a.no_traffic = a.no_terrain = a.no_c1 = True
        a.batch_scenes = 0
        a.guided_seconds = 0.0
    return a


def profile_wavefront(cr, L, scenes, tuning: str | None = None):
    """The wavefront integrator on the 1080p workload: one warm-up spp and 8
    timed spp (depth 16, MIS), for rocprofv3 kernel traces."""
    if tuning:
        cr.get_context(0).set_tuning(**json.loads(tuning))
    data, camd = scenes.spheres_scene()
    scene = cr.Scene(data)
    sensor = cr.Sensor(camd["size"])
    scene.setCamera(scenes.make_camera(camd, sensor))
    cfg = cr.Configs(maxDepth=16, rrThreshold=0.05, useMIS=True, rrStartDepth=3,
                     mode=L.MODE_WAVEFRONT)
    integ = cr.PathTraceIntegrator(cfg)
    integ.dispatch(scene, 1, readback=False)
    st = integ.dispatch(scene, 8, first_sample=1, readback=False)
    print(json.dumps({"spp": 8, "ms_per_spp": st["kernel_ms"] / 8}), flush=True)


def profile_guided(cr, L, scenes):
    """Launches of the fused guided megakernel (training and inference), the
    guided wavefront, the plain megakernel and the plain wavefront on the
    occluded room, 256 x 256, max depth 16, for rocprofv3 kernel traces and
    counter runs (MfmaUtil of the fused kernel)."""
    from craystal_amd.guiding import Guide, GuidedIntegrator
    data, camd = scenes.occluded_room()
    cfg = cr.Configs(maxDepth=16, rrThreshold=0.05, useMIS=True, rrStartDepth=3)
    scene = cr.Scene(data)
    sensor = cr.Sensor((256, 256))
    scene.setCamera(scenes.make_camera(camd, sensor))
    ctx = cr.get_context(0)
    P = data.concatenated()[0]
    guide = Guide(P.min(0), P.max(0), seed=11)
    gi = GuidedIntegrator(cfg, guide, alpha=0.5, lr=1e-2, train_batch=4096, train_steps=4,
                          materials=1, target=0)
    ctx.set_tuning(guided_fused=1)
    tr = gi.dispatch(scene, 16, first_sample=0, train=True, readback=False)
    inf = gi.dispatch(scene, 16, first_sample=16, train=False, readback=False)
    ctx.set_tuning(guided_fused=2)
    wfg = gi.dispatch(scene, 16, first_sample=16, train=False, readback=False)
    ctx.set_tuning(guided_fused=1)
    mk = cr.PathTraceIntegrator(cfg).dispatch(scene, 16, first_sample=32, readback=False)
    wf = cr.PathTraceIntegrator(cr.Configs(**{**cfg.__dict__, "mode": L.MODE_WAVEFRONT})).dispatch(
        scene, 16, first_sample=48, readback=False)
    print(json.dumps({"pixels": 65536, "spp_each": 16, "fused_guided_train_ms": tr["kernel_ms"],
                      "fused_guided_ms": inf["kernel_ms"], "wavefront_guided_ms": wfg["kernel_ms"],
                      "megakernel_ms": mk["kernel_ms"], "wavefront_ms": wf["kernel_ms"],
                      "guide_steps": guide.info()["steps"]}), flush=True)


def profile_kernels(cr, L, scenes, tuning: str | None = None):
    """Minimal launches for rocprofv3 --pmc runs (FETCH_SIZE / WRITE_SIZE): a
    256 MiB copy, the closest-hit batch counted and at the tuned shape, one
    megakernel launch (of as many sample indices as render_pool groups).
    tuning: JSON of crh_tuning fields to apply first (the launch shapes the
    timed run uses)."""
    if tuning:
        cr.get_context(0).set_tuning(**json.loads(tuning))
    nbytes = 256 << 20
    a, b = cr.DeviceBuffer(nbytes), cr.DeviceBuffer(nbytes)
    a.memset(1)
    b.copyFrom(a)
    data, camd = scenes.spheres_scene()
    scene = cr.Scene(data)
    n = 1 << 24
    rays = scenes.mixed_batch(data, camd, n)
    rb, hb = cr.DeviceBuffer(rays.nbytes), cr.DeviceBuffer(n * 16)
    rb.copyFromHost(rays)
    cnt = cr.trace_closest_device(scene, rb, n, hb, count_work=True)
    cr.trace_closest_device(scene, rb, n, hb)
    W, H = camd["size"]
    sensor = cr.Sensor((W, H))
    cam = scenes.make_camera(camd, sensor)
    scene.setCamera(cam)
    cfg = cr.Configs(maxDepth=camd["max_depth"], rrThreshold=0.05, useMIS=True,
                     rrStartDepth=camd["rr_start_depth"])
    # one megakernel launch of as many sample indices as the tuning renders together
    spl = max(1, cr.get_context(0).get_tuning()["render_pool"] // (W * H))
    st = cr.PathTraceIntegrator(cfg).dispatch(scene, spl, first_sample=0, readback=False)
    print(json.dumps({"copy_bytes": nbytes, "trace_rays": n,
                      "trace_node_visits": cnt["node_visits"], "trace_tri_tests": cnt["tri_tests"],
                      "render_samples": spl * W * H, "render_rays": st["rays"]}), flush=True)


def profile_terrain(cr, L, scenes, tuning: str | None = None):
    """Minimal launches on the 43.8M-triangle terrain (the fast GPU-built
    BVH of the timed leg) for rocprofv3: one spp of the four 3840 x 2160 views in one batched launch sequence (graph
    capture off), then one 16,777,216-ray closest-hit batch from the first
    view with work counting and one without."""
    from craystal_amd.batch import Batch
    ctx = cr.get_context(0)
    if tuning:
        ctx.set_tuning(**json.loads(tuning))
    ctx.set_tuning(batch_graph=2)
    data, cams = scenes.terrain_scene()
    data.bvhParams = scenes.TERRAIN_FAST_BVH   # the BVH the timed terrain leg uses
    data.bvhDevice = True
    scene = cr.Scene(data)
    batch = Batch([scene] * len(cams))
    sensors = [cr.Sensor(c["size"]) for c in cams]
    cam_objs = [scenes.make_camera(c, s) for c, s in zip(cams, sensors)]
    cfg = cr.Configs(maxDepth=cams[0]["max_depth"], rrThreshold=0.05, useMIS=True,
                     rrStartDepth=cams[0]["rr_start_depth"])
    st = batch.render(cam_objs, [1] * len(cams), cfg, first_sample=1, readback=False)
    n = 1 << 24
    rays = scenes.mixed_batch(data, cams[0], n)
    rb, hb = cr.DeviceBuffer(rays.nbytes), cr.DeviceBuffer(n * 16)
    rb.copyFromHost(rays)
    cnt = cr.trace_closest_device(scene, rb, n, hb, count_work=True)
    tr = cr.trace_closest_device(scene, rb, n, hb)
    print(json.dumps({"render_samples": 4 * cams[0]["size"][0] * cams[0]["size"][1],
                      "render_ms": round(st["kernel_ms"], 3), "render_rays": st["rays"],
                      "trace_rays": n, "trace_ms": round(tr["kernel_ms"], 3),
                      "trace_node_visits": cnt["node_visits"],
                      "trace_tri_tests": cnt["tri_tests"]}), flush=True)


def batch_leg(cr, scenes, n: int) -> dict:
    """The many-scene batch: n seeded scenes (triangles log-uniform in [1k,
    500k], resolutions 128^2..1024^2, 4/16/64 spp, max depth 12) rendered in
    one persistent launch per sample index, replayed from a captured graph,
    after the batch autotuner picked its tile shape, grid, ranges and stack.
    Also times the same scenes rendered one after another."""
    from craystal_amd import autotune
    from craystal_amd.batch import Batch
    items = [scenes.batch_scene(i) for i in range(n)]
    t0 = time.perf_counter()
    scene_list = [cr.Scene(d) for d, _ in items]
    build_s = time.perf_counter() - t0
    batch = Batch(scene_list)
    info = batch.info()
    sensors = [cr.Sensor(c["size"]) for _, c in items]
    cams = [scenes.make_camera(c, s) for (_, c), s in zip(items, sensors)]
    spp = [c["spp"] for _, c in items]
    cfg = cr.Configs(maxDepth=12, rrThreshold=0.05, useMIS=True, rrStartDepth=3)
    tuned = autotune.tune_batch(batch, cams, [1] * n, cfg, tile_ws=(8, 16, 32),
                                grids=(1024, 2048), chunks=(1, 8), stacks=(1, 3), waves=(4, 6, 8),
                                reps=1, slots=(1, 4, 16, 64), slot_spp=spp)
    for s in sensors:
        s.clear()
    batch.render(cams, spp, cfg, first_sample=0, readback=False)     # captures the graph
    for s in sensors:
        s.clear()
    st = batch.render(cams, spp, cfg, first_sample=0, readback=False)
    samples = sum(c["size"][0] * c["size"][1] * k for (_, c), k in zip(items, spp))
    # the batched wavefront (one launch per stage and bounce over every
    # scene's paths), its queue capacity and launch shapes tuned on 1 spp
    wtuned = autotune.tune_batch_wavefront(batch, cams, [1] * n, cfg)
    from craystal_amd import _lib as L
    wcfg = cr.Configs(**{**cfg.__dict__, "mode": L.MODE_WAVEFRONT})
    batch.render(cams, spp, wcfg, first_sample=0, readback=False)    # captures the graph
    wst = batch.render(cams, spp, wcfg, first_sample=0, readback=False)
    solo_ms = 0.0
    for sc, cam, k in zip(scene_list, cams, spp):
        sc.setCamera(cam)
        solo_ms += cr.PathTraceIntegrator(cfg).dispatch(sc, k, readback=False)["kernel_ms"]
    return {"scenes": n, "triangles": info["triangles"], "pool_mb": round(info["pool_bytes"] / 2**20, 1),
            "bvh_build_s": round(build_s, 2), "samples": samples,
            "kernel_ms": round(st["kernel_ms"], 3),
            "msamples_per_s": round(samples / (st["kernel_ms"] / 1e3) / 1e6, 2),
            "grays_per_s": round(st["rays"] / (st["kernel_ms"] / 1e3) / 1e9, 4),
            "solo_sequential_ms": round(solo_ms, 3),
            "wavefront": {"kernel_ms": round(wst["kernel_ms"], 3),
                          "msamples_per_s": round(samples / (wst["kernel_ms"] / 1e3) / 1e6, 2),
                          "tuning": {k: wtuned[k] for k in ("wf_pool", "wf_finish", "wf_stack",
                                                             "wf_trace_block", "wf_grid")}},
            "tuning": {k: tuned[k] for k in ("batch_tile_w", "batch_grid", "batch_chunks",
                                              "batch_stack", "batch_waves", "batch_slots")},
            "slots_ms": tuned["slots_ms"]}


def terrain_guided(cr, scene, data, cams, cfg, spp: int, px: float) -> dict:
    """Guided rendering of the terrain's four views in one fused launch per
    pass group (crh_render_guided_views) with the 128-wide guide: `spp`
    spp of every view with training (the guide learns from all views),
    then `spp` timed spp without.  One-sample MIS with alpha 0.5 at every
    Lambertian and Principled vertex."""
    from craystal_amd import autotune
    from craystal_amd.guiding import Guide, GuidedIntegrator
    P = data.concatenated()[0]
    guide = Guide(P.min(0), P.max(0), seed=17, width=128)
    gi = GuidedIntegrator(cfg, guide, alpha=0.5, lr=1e-2, train_batch=16384, train_steps=4,
                          materials=0, target=0)
    t0 = time.perf_counter()
    tr = gi.dispatch_views(scene, cams, spp, first_sample=1 << 20, train=True, readback=False)
    train_s = time.perf_counter() - t0
    for s in (c.getSensor() for c in cams):
        s.clear()
    t0 = time.perf_counter()
    st = gi.dispatch_views(scene, cams, spp, first_sample=1 << 21, train=False, readback=False)
    el = time.perf_counter() - t0
    n = px * spp
    return {"guide_width": 128, "spp_trained": spp, "spp_timed": spp,
            "adam_steps": guide.info()["steps"], "train_s": round(train_s, 3),
            "train_msamples_per_s": round(n / train_s / 1e6, 2),
            "kernel_ms": round(st["kernel_ms"], 3), "wall_ms": round(el * 1e3, 3),
            "msamples_per_s": round(n / el / 1e6, 2),
            "grays_per_s": round(st["rays"] / el / 1e9, 4),
            "tuning": {k: cr.get_context(0).get_tuning()[k]
                       for k in ("guided_stack", "guided_waves", "guided_grid", "guided_pool")}}


def terrain_leg(cr, scenes, spp: int, trace_rays: int, bvh: str = "fast",
                guided_spp: int = 2) -> dict:
    """The large-scale workload: the fBm terrain (4096 x 4096 vertices,
    33,538,050 triangles) with 2,000 subdivision-4 icospheres (43,778,066
    triangles with the sun disk), four 3840 x 2160 camera views rendered in
    one persistent launch per sample index (the scene held once in the pool),
    max depth 24, MIS.  One warm-up spp, then `spp` timed spp of every view.
    Also the closest-hit batch of `trace_rays` mixed rays from the first
    view, at the shape the autotuner picks for it.  bvh: "fast" builds the
    BVH on the GPU without the reference's depth cap of 20
    (scenes.TERRAIN_FAST_BVH), "reference" with the reference's parameters
    (leaves of up to 599 triangles)."""
    from craystal_amd import autotune
    from craystal_amd.batch import Batch
    t0 = time.perf_counter()
    data, cams = scenes.terrain_scene()
    gen_s = time.perf_counter() - t0
    progress(f"terrain generated: {data.triangle_count} triangles in {gen_s:.1f} s")
    if bvh == "fast":
        data.bvhParams = scenes.TERRAIN_FAST_BVH
        data.bvhDevice = True
    t0 = time.perf_counter()
    scene = cr.Scene(data)
    build_s = time.perf_counter() - t0
    bst = scene.bvh_stats()
    progress(f"terrain scene built in {build_s:.1f} s: {bst}")
    batch = Batch([scene] * len(cams))
    info = batch.info()
    sensors = [cr.Sensor(c["size"]) for c in cams]
    cam_objs = [scenes.make_camera(c, s) for c, s in zip(cams, sensors)]
    cfg = cr.Configs(maxDepth=cams[0]["max_depth"], rrThreshold=0.05, useMIS=True,
                     rrStartDepth=cams[0]["rr_start_depth"])
    batch.render(cam_objs, [1] * len(cams), cfg, first_sample=0, readback=False)
    t0 = time.perf_counter()
    st = batch.render(cam_objs, [spp] * len(cams), cfg, first_sample=1, readback=False)
    el = time.perf_counter() - t0
    W, H = cams[0]["size"]
    samples = W * H * spp * len(cams)
    out = {"scene": "fBm terrain 4096^2 vertices (8 octaves, 50 m over 1 km) + 2,000 "
                    "icospheres s4 + sun disk + sky",
           "triangles": int(data.triangle_count), "views": len(cams), "width": W, "height": H,
           "max_depth": cams[0]["max_depth"], "spp_timed": spp,
           "generate_s": round(gen_s, 2), "scene_create_s": round(build_s, 2),
           "bvh": {"params": "bins, max leaf, max depth = " + ", ".join(map(str, data.bvhParams)),
                   "built_on": "gpu" if data.bvhDevice else "host",
                   "nodes": bst["node_count"], "max_depth": bst["max_depth"],
                   "max_leaf": bst["max_leaf"], "build_s": round(bst["build_ms"] / 1e3, 3)},
           "pool_mb": round(info["pool_bytes"] / 2**20, 1),
           "kernel_ms": round(st["kernel_ms"], 3), "wall_ms": round(el * 1e3, 3),
           "msamples_per_s": round(samples / el / 1e6, 2),
           "grays_per_s": round(st["rays"] / el / 1e9, 4),
           "rays_per_sample": round(st["rays"] / samples, 4)}
    progress(f"terrain render: {out['msamples_per_s']} Msamples/s")
    if guided_spp > 0:
        out["guided"] = terrain_guided(cr, scene, data, cam_objs, cfg, guided_spp, samples / spp)
        progress(f"terrain guided: {out['guided']['msamples_per_s']} Msamples/s")
    del sensors, cam_objs, batch
    if trace_rays > 0:
        n = trace_rays
        rays = scenes.mixed_batch(data, cams[0], n)
        rb, hb = cr.DeviceBuffer(rays.nbytes), cr.DeviceBuffer(n * 16)
        rb.copyFromHost(rays)
        tt = autotune.tune_trace(scene, rays, reps=1, stacks=(2, 3), blocks=(256,),
                                 persistent=((2, 0), (1, 16), (1, 32)), pt_grids=(1024, 4096),
                                 tris=(1, 2))
        cr.trace_closest_device(scene, rb, n, hb)
        cnt = cr.trace_closest_device(scene, rb, n, hb, count_work=True)
        kms = float(np.mean([cr.trace_closest_device(scene, rb, n, hb)["kernel_ms"]
                             for _ in range(3)]))
        b_ray = B_RAY_FIXED + cnt["node_visits"] / n * B_NODE + cnt["tri_tests"] / n * B_TRI
        ach = b_ray * n / (kms / 1e3) / 1e9
        out["closest_hit"] = {
            "rays": n, "grays_per_s": round(n / (kms / 1e3) / 1e9, 4),
            "kernel_ms": round(kms, 3),
            "node_visits_per_ray": round(cnt["node_visits"] / n, 3),
            "tri_tests_per_ray": round(cnt["tri_tests"] / n, 3),
            "tuning": {k: tt[k] for k in ("stack", "trace_block", "persistent", "refill",
                                          "pt_grid", "trace_tris")},
            "roofline": {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 5),
                         "bytes_per_ray": round(b_ray, 1)}}
        progress(f"terrain closest hit: {out['closest_hit']['grays_per_s']} Grays/s")
    return out


def c1_leg(cr, L, scenes, with_cpu: bool, reps: int = 20) -> dict:
    """The small plain configuration: the spectral Cornell box (36 triangles,
    smooth 95-bin reflectance spectra, constant-spectrum emitter), 256 x 256,
    16 spp, max depth 8, roulette from depth 3, MIS, hero-wavelength transport
    into a CIE XYZ film.  The megakernel and the wavefront are tuned on it
    (every candidate gives the same film), then each renders the 16 spp
    `reps` times; the mean rate is reported.  The CPU oracle renders the same
    16 spp on all host cores and on one."""
    from craystal_amd import autotune
    data, camd = scenes.cornell_box_spectral()
    W, H = 256, 256
    scene = cr.Scene(data)
    sensor = cr.Sensor((W, H))
    cam = scenes.make_camera(camd, sensor)
    scene.setCamera(cam)
    cfg = cr.Configs(maxDepth=8, rrThreshold=0.05, useMIS=True, rrStartDepth=3, spectral=True)
    spp = 16
    tuned = autotune.tune_render(scene, cfg, spp=spp, stacks=(1, 2, 3), render_waves=(4, 6, 8),
                                 render_regen=((2, 0), (1, 1024), (1, 2048)),
                                 render_pools=(1 << 20, 1 << 21), wf_grids=(4096, 16384),
                                 wf_blocks=(128, 256), wf_persistent=((2, 0),),
                                 wf_pools=(1 << 20, 1 << 21), wf_finishes=(65, 1, 2))
    samples = W * H * spp
    out = {"scene": "spectral Cornell box, 36 tris, 256x256, 16 spp, max depth 8, RR from "
                    "depth 3, MIS, 95-bin spectra (hero wavelength), CIE XYZ film",
           "samples_per_render": samples}
    for name, mode in (("megakernel", L.MODE_MEGAKERNEL), ("wavefront", L.MODE_WAVEFRONT)):
        integ = cr.PathTraceIntegrator(cr.Configs(**{**cfg.__dict__, "mode": mode}))
        integ.dispatch(scene, spp, first_sample=0, readback=False)
        t0 = time.perf_counter()
        for r in range(reps):
            integ.dispatch(scene, spp, first_sample=(r + 1) * spp, readback=False)
        el = time.perf_counter() - t0
        out[name] = {"msamples_per_s": round(samples * reps / el / 1e6, 2),
                     "ms_per_render": round(el / reps * 1e3, 3)}
    out["tuning"] = {k: tuned[k] for k in ("mode", "render_stack", "render_waves",
                                            "render_regen", "render_pool", "wf_pool",
                                            "wf_finish")}
    if with_cpu:
        import oracle
        orc = oracle.Scene(data)
        threads = int(oracle.lib().orc_max_threads())
        t0 = time.perf_counter()
        orc.render(cam, (W, H), cfg, 0, spp)
        c = time.perf_counter() - t0
        t0 = time.perf_counter()
        orc.render(cam, (W, H), cfg, 0, 2, threads=1)
        c1 = time.perf_counter() - t0
        out["cpu_oracle"] = {"msamples_per_s": round(samples / c / 1e6, 3), "cores": threads,
                             "seconds": round(c, 3), "cpu_model": _cpu_model(),
                             "one_core_msamples_per_s": round(W * H * 2 / c1 / 1e6, 3),
                             "one_core_sample": "2 of the 16 spp on one thread"}
    return out


def guided_leg(cr, scenes, seconds: float, repeats: int, size: int = 1024,
               materials: int = 1, train_steps: int = 32, train_batch: int = 16384) -> dict:
    """Guided vs plain on the occluded-light room (200 objects, 171,388
    triangles, max depth 16) at size x size against the committed
    65,536-spp plain image of the same size: equal-time relMSE and the time
    to relMSE 0.01.  The reference's estimator is heavy-tailed (rare, very
    bright samples from its Principled BSDF and the emitter double count),
    so one run's relMSE at a given time varies by several times between
    independent runs; each method therefore renders `repeats` independent
    runs (disjoint sample indices, a fresh guide per run) of `seconds`
    render time and the medians are reported.  Where a run does not reach
    0.01 within its time, its time to 0.01 is extrapolated from relMSE ~ 1/t
    fitted on the second half of its curve.  Guided: the fused guided
    megakernel (guide inference on MFMA inside the path kernel), one-sample
    MIS with alpha 0.5 at the vertices whose BSDF `materials` selects (1:
    Lambertian, 2: Principled, 3: both), the guide training during the first
    25 % of each run.  The plain megakernel's launch shape is tuned on this
    scene first, so both methods run at their best."""
    from craystal_amd import autotune
    from craystal_amd import experiments as X
    ref_path = ROOT / "craystal_amd" / "data" / f"room{size}_plain65536.npy"
    ref = np.load(ref_path).astype(np.float32)
    data, camd = scenes.occluded_room()
    cfg = cr.Configs(maxDepth=16, rrThreshold=0.05, useMIS=True, rrStartDepth=3)
    scene = cr.Scene(data)
    sensor = cr.Sensor((size, size))
    scene.setCamera(scenes.make_camera(camd, sensor))
    px = size * size
    tuned = autotune.tune_render(scene, cfg, spp=4 if size > 256 else 16, stacks=(2, 3),
                                 render_waves=(6, 8),
                                 render_regen=((2, 0), (1, 1024), (1, 2048), (1, 4096)),
                                 wf_grids=(16384,), wf_blocks=(128,), wf_persistent=((2, 0),),
                                 wf_finishes=(65,))
    # the fused guided kernel's launch shape, tuned with a guide trained for
    # a few passes (training changes which vertices are guided and how long
    # paths get; the tuning does not change any image)
    from craystal_amd.guiding import Guide, GuidedIntegrator
    P = data.concatenated()[0]
    gi = GuidedIntegrator(cfg, Guide(P.min(0), P.max(0), seed=5), alpha=0.5, lr=1e-2,
                          train_batch=4096, train_steps=4, materials=materials, target=0)
    for k in range(4):
        gi.dispatch(scene, 4, first_sample=(1 << 26) + 4 * k, train=True, readback=False)
    gtuned = autotune.tune_guided(gi, scene, spp=4 if size > 256 else 16, waves=(6, 8),
                                  grids=(2048, 4096, 8192), pools=(None, 1 << 24))
    del gi
    sensor.clear()
    chunk = 16 if size > 256 else 64
    runs = {"plain": [], "guided": []}
    for r in range(repeats):
        first = (r + 1) << 21
        pc = X.plain_curve(scene, sensor, cfg, ref, seconds, chunk, first=first)
        runs["plain"].append(X.summarize(pc, 0.01, px))
        gc, info = X.guided_curve(scene, sensor, data, cfg, ref, seconds, chunk, first=first,
                                  seed=11 + r, materials=materials, target=0,
                                  train_batch=train_batch, train_steps=train_steps, lr=1e-2,
                                  alpha=0.5)
        runs["guided"].append({**X.summarize(gc, 0.01, px), **info})
        progress(f"guided leg run {r}: relMSE plain {runs['plain'][-1]['final_relmse']:.4f} "
                 f"guided {runs['guided'][-1]['final_relmse']:.4f}")
    out = {"scene": f"occluded room, 200 objects, 171,388 tris, {size}x{size}, max depth 16",
           "target_relmse": 0.01,
           "plain_tuning": {k: tuned[k] for k in ("render_stack", "render_waves", "render_regen",
                                                        "render_grid")},
           "guided_kernel": "fused guided megakernel (kRenderGuided)",
           "guided_materials": {1: "Lambertian", 2: "Principled", 3: "Lambertian and Principled"}[materials],
           "train_steps_per_pass": train_steps, "train_batch": train_batch,
           "guided_tuning": {k: gtuned[k] for k in ("guided_stack", "guided_waves",
                                                     "guided_grid", "guided_pool")},
           "seconds_per_run": seconds, "runs": repeats,
           "reference": f"craystal_amd/data/{ref_path.name} (plain megakernel, 65,536 spp)"}
    for k, v in runs.items():
        out[k] = {"relmse_at_end": [round(x["final_relmse"], 5) for x in v],
                  "median_relmse_at_end": float(np.median([x["final_relmse"] for x in v])),
                  "extrapolated_seconds": [round(x["extrapolated_seconds"], 2) for x in v],
                  "median_extrapolated_seconds": float(np.median([x["extrapolated_seconds"]
                                                                 for x in v])),
                  "msamples_per_s": round(float(np.mean([x["msamples_per_s"] for x in v])), 2),
                  "spp_at_end": [x["final_spp"] for x in v],
                  "reached_target_s": [x["seconds_to_target"] for x in v],
                  "relmse_trimmed_at_end": [round(x["final_relmse_trimmed"], 5) for x in v],
                  "median_relmse_trimmed_at_end": float(np.median(
                      [x["final_relmse_trimmed"] for x in v]))}
    out["guided"]["adam_steps_per_run"] = [x["train_steps"] for x in runs["guided"]]
    out["guided"]["train_seconds_per_run"] = [x["train_seconds"] for x in runs["guided"]]
    out["speedup_median_time_to_target"] = round(
        out["plain"]["median_extrapolated_seconds"] / out["guided"]["median_extrapolated_seconds"],
        3)
    out["equal_time_relmse_ratio"] = round(
        out["plain"]["median_relmse_at_end"] / out["guided"]["median_relmse_at_end"], 3)
    # diagnostic: the same ratio on relMSE without the 0.1 % of pixels with the
    # largest error (equal render times, relMSE ~ 1/t)
    out["speedup_trimmed_diagnostic"] = round(
        out["plain"]["median_relmse_trimmed_at_end"] / out["guided"]["median_relmse_trimmed_at_end"],
        3)
    out["guided_vs_plain_sample_rate"] = round(
        out["guided"]["msamples_per_s"] / out["plain"]["msamples_per_s"], 3)
    return out


def measure_traffic(tuning: dict | None = None) -> dict | None:
    """HBM bytes per dispatch of the profiled kernels from two rocprofv3 --pmc
    runs of `bench.py --profile-kernels` (FETCH_SIZE, then WRITE_SIZE, KiB per
    dispatch), each scaled by what the counter reads for the 256 MiB copy
    that mode starts with.  Run as child processes before this process
    touches the GPU; None when rocprofv3 is missing or a run fails."""
    import csv
    import shutil
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None
    vals: dict[str, dict[str, float]] = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="crh_pmc_", dir="/tmp")
        cmd = ["timeout", "-s", "KILL", "240", exe, "--pmc", ctr, "--kernel-trace",
               "--output-format", "csv", "-d", d, "-o", "run", "--", sys.executable,
               str(ROOT / "bench.py"), "--profile-kernels"]
        if tuning:
            cmd += ["--profile-tuning", json.dumps(tuning)]
        r = subprocess.run(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"),
                           capture_output=True, text=True)
        files = list(Path(d).rglob("*counter_collection.csv"))
        if r.returncode != 0 or not files:
            return None
        for row in csv.DictReader(open(files[0])):
            name = row["Kernel_Name"]
            if "kRenderPass<" in name or "kRenderRegen<" in name:
                vals.setdefault("render_kernel", {})["name"] = name.split("<")[0]
            key = ("copy" if "copyBuffer" in name else
                   "render" if ("kRenderPass<" in name or "kRenderRegen<" in name) else
                   "kTraceClosest" if ("kTraceClosest<false" in name or
                                       "kTracePersistent<false" in name) else None)
            if key:
                # the calibration copy is the largest copy dispatch
                v = float(row["Counter_Value"])
                d = vals.setdefault(key, {})
                d[ctr] = max(d.get(ctr, 0.0), v) if key == "copy" else v
        shutil.rmtree(d, ignore_errors=True)
    try:
        copy_kib = 256 * 1024
        fs = copy_kib / vals["copy"]["FETCH_SIZE"]
        ws = copy_kib / vals["copy"]["WRITE_SIZE"]
        out = {"fetch_scale": round(fs, 5), "write_scale": round(ws, 5),
               "render_kernel": vals["render_kernel"]["name"]}
        for k in ("render", "kTraceClosest"):
            out[k] = int((vals[k]["FETCH_SIZE"] * fs + vals[k]["WRITE_SIZE"] * ws) * 1024)
        return out
    except (KeyError, ZeroDivisionError):
        return None


def main():
    a = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print(json.dumps({"error": "this benchmark measures one GPU only; multi-GPU runs are "
                                   "not part of it (use --gpus 1)"}), flush=True)
        sys.exit(2)

    traffic = None

    import craystal_amd as cr
    from craystal_amd import _lib as L
    from craystal_amd import scenes

    cr.load()
    if cr.device_count() < 1:
        raise SystemExit("no GPU visible to libcrayhip.so")
    if a.profile_kernels:
        profile_kernels(cr, L, scenes, a.profile_tuning)
        return
    if a.profile_guided:
        profile_guided(cr, L, scenes)
        return
    if a.profile_wavefront:
        profile_wavefront(cr, L, scenes, a.profile_tuning)
        return
    if a.profile_terrain:
        profile_terrain(cr, L, scenes, a.profile_tuning)
        return

    data, camd = scenes.spheres_scene()
    W, H = camd["size"]
    t0 = time.perf_counter()
    scene = cr.Scene(data)
    build_s = time.perf_counter() - t0
    bst = scene.bvh_stats()
    sensor = cr.Sensor((W, H))
    cam = scenes.make_camera(camd, sensor)
    scene.setCamera(cam)
    cfg = cr.Configs(maxDepth=camd["max_depth"], rrThreshold=0.05, useMIS=True,
                     rrStartDepth=camd["rr_start_depth"])
    integ = cr.PathTraceIntegrator(cfg)

    # launch shapes picked by the autotuner (outside the timed region; every
    # candidate must render the same image bit for bit)
    from craystal_amd import autotune
    if a.tuning:
        # launch shapes given (a previous run's "tuning"): no tuner launches,
        # so a profiler sees only the timed run's kernels
        ctx0 = cr.get_context(0)
        ctx0.set_tuning(**{k: v for k, v in json.loads(a.tuning).items()
                           if k in ctx0.TUNING_FIELDS})
        t0_ = ctx0.get_tuning()
        tuned = {**t0_, "wf_grid": t0_["wf_shade_grid"], "megakernel_ms": {},
                 "wavefront_ms": {}, "render_pool_ms": {}}
    else:
        tuned = autotune.tune_render(
            scene, cfg, spp=16, stacks=(1, 2, 3), wf_grids=(16384,), wf_blocks=(128, 256),
            wf_persistent=((2, 0), (1, 16, 2048), (1, 16, 4096)),
            wf_pools=(1 << 23, 1 << 24, 1 << 25), render_waves=(1, 4, 6, 8),
            render_regen=((2, 0), (1, 1024), (1, 2048), (1, 4096)),
            wf_finishes=(65, 1, 2, 3), render_bands=(1, 2),
            render_pools=(1 << 21, 1 << 23, 1 << 24, 1 << 25, 1 << 26), pool_spp=32)
    sensor.clear()

    progress(f"scene built, autotuned: {tuned['render_stack']=} {tuned['wf_pool']=}")
    # launch shapes the HBM-traffic child runs use (the timed ones)
    prof_tuning = {k: tuned[k] for k in ("render_stack", "render_waves", "render_regen",
                                         "render_band", "render_pool")}
    if tuned["render_regen"] == 1:
        prof_tuning["render_grid"] = tuned["render_grid"]
    # warmup (untimed)
    if a.warmup > 0:
        integ.dispatch(scene, a.warmup, first_sample=0, readback=False)
    # work counters for the roofline, outside the timed region
    cnt_cfg = cr.Configs(**{**cfg.__dict__})
    cc = cnt_cfg.to_c()
    cc.count_work = 1
    import ctypes as C
    st = L.crh_stats()
    L.call("crh_render", scene.handle, C.byref(cam.data), C.byref(cc), 100000, 1,
           sensor.handle, C.byref(st))
    px = W * H
    rays_ps = st.rays / px
    nodes_ps = st.node_visits / px
    tris_ps = st.tri_tests / px
    b_sample = rays_ps * B_RAY_FIXED + nodes_ps * B_NODE + tris_ps * B_TRI + B_FILM
    sensor.clear()

    # timed region: K steps = K spp passes of the full image
    t0 = time.perf_counter()
    stats = integ.dispatch(scene, a.steps, first_sample=a.warmup, readback=False)
    elapsed = time.perf_counter() - t0
    samples = px * a.steps
    msps = samples / elapsed / 1e6
    main_ms = stats["kernel_ms"]
    # dominant kernel: the megakernel launches of the timed run, timed with
    # HIP events on their stream (launches = pass groups; the stats count
    # them plus one film update per spp)
    launches = max(1, int(stats["launches"]) - a.steps)
    avg_launch_s = stats["main_kernel_ms"] / launches / 1e3
    samples_per_launch = a.steps * px / launches
    achieved = b_sample * samples_per_launch / avg_launch_s / 1e9

    out = {
        "metric": "Msamples/s + time-to-relMSE 0.01, guided vs plain, 1 MI355X; Grays/s",
        "value": round(msps, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": "spheres-1M: 48 icospheres s5 + ground + area light + env, "
                               "983,044 tris, 1920x1080, 1 spp per step, max depth 16, "
                               "RR from depth 3, MIS, plain megakernel",
                   "triangles": int(scene.triangle_count), "width": W, "height": H,
                   "spp_timed": a.steps},
        "grays_per_s": round(stats["rays"] / elapsed / 1e9, 4),
        "rays_per_sample": round(rays_ps, 4),
        "gpu_kernel_ms": round(main_ms, 3),
        "bvh": {"nodes": bst["node_count"], "max_depth": bst["max_depth"],
                "max_leaf": bst["max_leaf"], "build_s": round(build_s, 3)},
        "roofline": {"bound": "hbm",
                     "kernel": "kRenderRegen" if tuned["render_regen"] == 1 else "kRenderPass",
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": traffic["render"] if traffic else None,
                     "bytes_per_sample": round(b_sample, 1),
                     "avg_launch_ms": round(avg_launch_s * 1e3, 4),
                     "samples_per_launch": int(samples_per_launch),
                     "node_visits_per_sample": round(nodes_ps, 3),
                     "tri_tests_per_sample": round(tris_ps, 3)},
    }

    progress(f"megakernel: {msps:.1f} Msamples/s")
    # wavefront integrator on the same workload and timed steps
    wf_cfg = cr.Configs(**{**cfg.__dict__, "mode": L.MODE_WAVEFRONT})
    wf = cr.PathTraceIntegrator(wf_cfg)
    wf.dispatch(scene, 1, first_sample=0, readback=False)
    t0 = time.perf_counter()
    wst = wf.dispatch(scene, a.steps, first_sample=a.warmup, readback=False)
    wel = time.perf_counter() - t0
    out["tuning"] = {k: tuned[k] for k in ("render_stack", "render_waves", "render_regen",
                                           "render_grid", "render_band", "render_pool",
                                           "wf_stack", "wf_grid",
                                           "wf_trace_block", "wf_trace_grid", "wf_persistent",
                                           "wf_refill",
                                           "wf_pool", "wf_finish")}
    out["tuning"]["wavefront_ms_per_spp"] = tuned["wavefront_ms"]
    out["tuning"]["megakernel_ms_per_spp"] = tuned["megakernel_ms"]
    out["tuning"]["render_pool_ms_per_spp"] = tuned["render_pool_ms"]
    out["wavefront"] = {"msamples_per_s": round(samples / wel / 1e6, 3),
                        "ms_per_step": round(wel / a.steps * 1e3, 3),
                        "grays_per_s": round(wst["rays"] / wel / 1e9, 4)}

    progress(f"wavefront: {out['wavefront']['msamples_per_s']} Msamples/s")
    rays = None
    if not a.no_trace:
        n = a.trace_rays
        rays = scenes.mixed_batch(data, camd, n)
        rb = cr.DeviceBuffer(rays.nbytes)
        hb = cr.DeviceBuffer(n * 16)
        rb.copyFromHost(rays)
        if a.tuning:
            ttuned = cr.get_context(0).get_tuning()
        else:
            ttuned = autotune.tune_trace(scene, rays, reps=2, stacks=(2, 3),
                                         blocks=(64, 128, 256),
                                         persistent=((2, 0), (1, 16), (1, 32)),
                                         pt_grids=(2048, 4096), tris=(1, 2))
        prof_tuning.update(stack=ttuned["stack"], trace_block=ttuned["trace_block"],
                           persistent=ttuned["persistent"])
        if ttuned["persistent"] == 1:
            prof_tuning.update(refill=ttuned["refill"], pt_grid=ttuned["pt_grid"],
                               trace_tris=ttuned["trace_tris"])
        progress(f"profile tuning: {json.dumps(prof_tuning)}")
        if not a.no_traffic:
            # HBM traffic of both kernels at the launch shapes chosen, from two
            # rocprofv3 child runs while this process is idle
            traffic = measure_traffic(prof_tuning)
            progress(f"traffic counters: {traffic}")
            if traffic:
                out["roofline"]["traffic"] = traffic["render"]
        cr.trace_closest_device(scene, rb, n, hb)  # warmup
        cnt = cr.trace_closest_device(scene, rb, n, hb, count_work=True)
        times = []
        for _ in range(a.trace_reps):
            times.append(cr.trace_closest_device(scene, rb, n, hb)["kernel_ms"])
        kms = float(np.mean(times))
        b_ray = B_RAY_FIXED + cnt["node_visits"] / n * B_NODE + cnt["tri_tests"] / n * B_TRI
        ach = b_ray * n / (kms / 1e3) / 1e9
        out["closest_hit"] = {
            "rays": n, "grays_per_s": round(n / (kms / 1e3) / 1e9, 4),
            "kernel_ms": round(kms, 3),
            "kernel": "kTracePersistent" if ttuned["persistent"] == 1 else "kTraceClosest",
            "node_visits_per_ray": round(cnt["node_visits"] / n, 3),
            "tri_tests_per_ray": round(cnt["tri_tests"] / n, 3),
            "tuning": {k: ttuned[k] for k in ("stack", "trace_block", "persistent", "refill",
                                              "pt_grid", "trace_tris")},
            "roofline": {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 5),
                         "traffic": traffic["kTraceClosest"] if traffic else None,
                         "bytes_per_ray": round(b_ray, 1)}}
        del rb, hb

    if a.no_trace and not a.no_traffic:
        traffic = measure_traffic(prof_tuning)
        if traffic:
            out["roofline"]["traffic"] = traffic["render"]
    if traffic:
        out["roofline"]["traffic_note"] = (
            "HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of "
            "bench.py --profile-kernels (one 1920x1080 pass of "
            f"{traffic['render_kernel']} at the tuned launch shape; the closest-hit batch "
            "likewise), scaled by a 256 MiB copy "
            f"(fetch x {traffic['fetch_scale']}, write x {traffic['write_scale']})")
    if "closest_hit" in out:
        progress(f"closest hit: {out['closest_hit']['grays_per_s']} Grays/s")
    # Each further leg tunes launch shapes for its own workload; every one
    # starts from the tuning the main workload ran with.
    ctx = cr.get_context(0)
    main_tuning = ctx.get_tuning()
    if not a.no_c1:
        ctx.set_tuning(**main_tuning)
        out["c1"] = c1_leg(cr, L, scenes, not a.no_cpu)
        progress(f"C1: megakernel {out['c1']['megakernel']['msamples_per_s']}, wavefront "
                 f"{out['c1']['wavefront']['msamples_per_s']} Msamples/s")
    batch_tuning = {}
    if a.batch_scenes > 0:
        ctx.set_tuning(**main_tuning)
        out["batch"] = batch_leg(cr, scenes, a.batch_scenes)
        batch_tuning = out["batch"]["tuning"]
        progress(f"batch: {out['batch']['msamples_per_s']} Msamples/s")
    if a.guided_seconds > 0 and a.guided_runs > 0:
        ctx.set_tuning(**main_tuning)
        out["guided"] = guided_leg(cr, scenes, a.guided_seconds, a.guided_runs, a.guided_size,
                                   a.guided_materials, a.guided_train_steps,
                                   a.guided_train_batch)
    if not a.no_terrain:
        # the four views render as a batch: at the batch leg's tuned shape
        ctx.set_tuning(**main_tuning)
        ctx.set_tuning(**batch_tuning)
        out["terrain"] = terrain_leg(cr, scenes, a.terrain_spp, a.trace_rays, a.terrain_bvh,
                                     a.terrain_guided_spp)

    if not a.no_cpu:
        import oracle
        orc_scene = oracle.Scene(data)
        threads = int(oracle.lib().orc_max_threads())
        # spp passes of the full image until at least 10 s of CPU work
        passes, crays, t0 = 0, 0, time.perf_counter()
        while passes == 0 or time.perf_counter() - t0 < 10.0:
            _, _, r = orc_scene.render_samples(cam, (W, H), cfg, passes)
            crays += r
            passes += 1
        cs = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(passes * px / cs / 1e6, 4), "unit": "Msamples/s",
                               "cores": threads, "kind": "port", "cpu_model": _cpu_model(),
                               "sample": f"{passes} spp passes of the same 1920x1080 image "
                                         f"({passes * px} paths, {crays} rays) by the C++ "
                                         f"oracle, OpenMP over rows",
                               "seconds": round(cs, 3),
                               "grays_per_s": round(crays / cs / 1e9, 5)}
        # one core: one spp pass of the same image
        t0 = time.perf_counter()
        _, _, r1 = orc_scene.render_samples(cam, (W, H), cfg, passes, threads=1)
        c1 = time.perf_counter() - t0
        out["cpu_baseline"]["one_core"] = {
            "value": round(px / c1 / 1e6, 4), "unit": "Msamples/s", "seconds": round(c1, 3),
            "sample": "one spp pass of the same 1920x1080 image on one thread",
            "grays_per_s": round(r1 / c1 / 1e9, 5)}
        if rays is not None:
            # the closest-hit batch in full on all cores, and its first 2^20 rays on one
            t0 = time.perf_counter()
            orc_scene.trace_closest(rays)
            ct = time.perf_counter() - t0
            m = min(len(rays), 1 << 20)
            t0 = time.perf_counter()
            orc_scene.trace_closest(rays[:m], threads=1)
            ct1 = time.perf_counter() - t0
            out["cpu_baseline"]["closest_hit"] = {
                "grays_per_s": round(len(rays) / ct / 1e9, 5), "seconds": round(ct, 3),
                "cores": threads, "rays": len(rays),
                "one_core_grays_per_s": round(m / ct1 / 1e9, 5), "one_core_rays": m}
        progress(f"cpu baseline: {out['cpu_baseline']['value']} Msamples/s")
    print(json.dumps(out), flush=True)


def _cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo, as lscpu reports it)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
