"""Autotuner for launch shapes that do not change results.

Picks, by short timed trials on the device, the traversal-stack placement
(LDS or private memory), threads per block of the ray-batch kernels, and the
grid and block sizes of the wavefront kernels (its queue work distribution).
Every candidate renders / traces the same inputs, and the tuner checks that
all candidates produce bit-identical outputs before applying the fastest.
"""
from __future__ import annotations

import itertools
import time

import numpy as np

from . import _lib as L
from .core import (Configs, DeviceBuffer, PathTraceIntegrator, Scene, get_context,
                   trace_closest_device)


def _median(xs):
    return float(np.median(np.asarray(xs)))


def tune_trace(scene: Scene, rays: np.ndarray, reps: int = 3, device: int = 0,
               stacks=(1, 2), blocks=(64, 128, 256), persistent=((2, 0),),
               pt_grids=(1024,), tris=(1,)) -> dict:
    """Tune the closest-hit batch kernel over stack placement, block size and
    persistence ((mode, refill) pairs; mode 1 = lanes refill, 2 = off; with
    refill, also the persistent grid and the triangles per traversal step);
    returns the chosen tuning and the per-candidate timings (ms)."""
    ctx = get_context(device)
    saved = ctx.get_tuning()
    n = len(rays)
    rb = DeviceBuffer(rays.nbytes, device)
    hb = DeviceBuffer(n * 16, device)
    rb.copyFromHost(np.ascontiguousarray(rays, dtype=L.RAY_DTYPE))
    ref = None
    results = {}
    cands = []
    for st, bl, (pm, rf) in itertools.product(stacks, blocks, persistent):
        for pg in (pt_grids if pm == 1 else (0,)):
            for tr in (tris if pm == 1 else (1,)):
                cands.append((st, bl, pm, rf, pg, tr))

    def kw_of(st, bl, pm, rf, pg, tr):
        kw = dict(stack=st, trace_block=bl, persistent=pm)
        if pm == 1:
            kw.update(refill=rf, pt_grid=pg, trace_tris=tr)
        return kw

    for c in cands:
        kw = kw_of(*c)
        ctx.set_tuning(**kw)
        trace_closest_device(scene, rb, n, hb)
        ms = [trace_closest_device(scene, rb, n, hb)["kernel_ms"] for _ in range(reps)]
        out = np.zeros(n, L.HIT_DTYPE)
        hb.copyToHost(out)
        if ref is None:
            ref = out
        elif out.tobytes() != ref.tobytes():
            ctx.set_tuning(**saved)
            raise AssertionError(f"tuning {kw} changed the hits")
        results[c] = _median(ms)
    best = min(results, key=results.get)
    st, bl, pm, rf, pg, tr = best
    ctx.set_tuning(**kw_of(*best))
    return {"stack": st, "trace_block": bl, "persistent": pm, "refill": rf, "pt_grid": pg,
            "trace_tris": tr,
            "timings_ms": {f"stack={k[0]},block={k[1]},persistent={k[2]},refill={k[3]},"
                           f"grid={k[4]},tris={k[5]}": round(v, 4) for k, v in results.items()}}


def tune_render(scene: Scene, cfg: Configs, spp: int = 1, device: int = 0,
                stacks=(1, 2), wf_grids=(1024, 2048, 4096), wf_blocks=(128, 256),
                wf_persistent=((2, 0), (1, 16)), wf_pools=(None,), render_waves=(1,),
                render_regen=((2, 0), (1, 1024), (1, 2048)), wf_finishes=(65,),
                render_bands=(1,), render_pools=(None,), pool_spp: int | None = None,
                first_sample: int = 1 << 20) -> dict:
    """Tune the megakernel stack placement, register budget and path
    regeneration ((mode, persistent grid) pairs, mode 2 = one path per
    thread), then, for the best of those, its paths per launch
    (render_pools, timed over pool_spp samples, default spp; None keeps the
    current value), and the wavefront stack
    placement,
    grid/block sizes, persistence ((mode, refill) or (mode, refill, traversal
    grid) tuples; mode 1 = lanes refill, 2 = off), path pool (paths in
    flight; None keeps the current one) and finishing depth on the scene's current
    camera/sensor; every candidate must give the same image bit for bit.
    Applies the best choice of each integrator and returns the faster
    integrator's mode with the timings (ms per spp) of both."""
    ctx = get_context(device)
    sensor = scene.getCamera().getSensor()
    res_mk, res_wf = {}, {}
    ref = None

    def run(c, n=spp):
        sensor.clear()
        st = PathTraceIntegrator(c).dispatch(scene, n, first_sample=first_sample)
        return st["kernel_ms"] / n, sensor.data.copy()

    def mk_kw(st, wv, rg, grid, band):
        return dict(render_stack=st, render_waves=wv, render_regen=rg, render_band=band,
                    **({"render_grid": grid} if rg == 1 else {}))

    # (regeneration, grid, band): the XCD banding applies to one path per thread
    mk_modes = [(rg, grid, band) for rg, grid in render_regen
                for band in (render_bands if rg == 2 else render_bands[:1])]
    for st, wv, (rg, grid, band) in itertools.product(stacks, render_waves, mk_modes):
        ctx.set_tuning(**mk_kw(st, wv, rg, grid, band))
        run(cfg)
        ms, img = run(cfg)
        ref = img if ref is None else ref
        assert img.tobytes() == ref.tobytes(), "megakernel launch shape changed the image"
        res_mk[(st, wv, rg, grid, band)] = ms
    best_mk = min(res_mk, key=res_mk.get)
    best_stack, best_waves, best_regen, best_grid, best_band = best_mk
    ctx.set_tuning(**mk_kw(*best_mk))
    # paths per launch (sample indices rendered together) for the chosen shape
    rpool0 = ctx.get_tuning()["render_pool"]
    # Over the same spp the first stage's image is the reference; over other
    # spp the stage's own first image is (and the stage-1 image is checked
    # again below, once the pool is chosen).
    same_spp = pool_spp is None or pool_spp == spp
    res_pool, ref_pool = {}, (ref if same_spp else None)
    for p in [rpool0 if p is None else p for p in render_pools]:
        ctx.set_tuning(render_pool=p)
        run(cfg, pool_spp or spp)
        ms, img = run(cfg, pool_spp or spp)
        ref_pool = img if ref_pool is None else ref_pool
        assert img.tobytes() == ref_pool.tobytes(), "megakernel launch shape changed the image"
        res_pool[p] = ms
    best_pool = min(res_pool, key=res_pool.get)
    ctx.set_tuning(render_pool=best_pool)
    # The megakernel as it now stands (chosen shape and pool), timed over spp
    # like the wavefront below, so the mode choice compares like with like.
    if same_spp:
        mk_ms = min(res_mk[best_mk], res_pool[best_pool])
    else:
        run(cfg)
        mk_ms, img = run(cfg)
        assert img.tobytes() == ref.tobytes(), "megakernel launch shape changed the image"
    wf_cfg = Configs(**{**cfg.__dict__, "mode": L.MODE_WAVEFRONT})
    pool0 = ctx.get_tuning()["wf_pool"]
    pools = [pool0 if p is None else p for p in wf_pools]

    def wf_kw(st, g, b, pm, rf, tg, pool, fin):
        return dict(wf_stack=st, wf_trace_grid=tg or g, wf_shade_grid=g, wf_trace_block=b,
                    wf_persistent=pm, wf_pool=pool, wf_finish=fin,
                    **({"wf_refill": rf} if pm == 1 else {}))

    pers = [tuple(p) + (0,) * (3 - len(p)) for p in wf_persistent]
    for st, g, b, (pm, rf, tg), pool, fin in itertools.product(stacks, wf_grids, wf_blocks,
                                                               pers, pools, wf_finishes):
        ctx.set_tuning(**wf_kw(st, g, b, pm, rf, tg, pool, fin))
        ms, img = run(wf_cfg)
        assert img.tobytes() == ref.tobytes(), "wavefront launch shape changed the image"
        res_wf[(st, g, b, pm, rf, tg, pool, fin)] = ms
    best_wf = min(res_wf, key=res_wf.get)
    st, g, b, pm, rf, tg, pool, fin = best_wf
    ctx.set_tuning(**wf_kw(st, g, b, pm, rf, tg, pool, fin))
    wf_ms = res_wf[best_wf]
    return {"mode": L.MODE_WAVEFRONT if wf_ms < mk_ms else L.MODE_MEGAKERNEL,
            "render_stack": best_stack, "render_waves": best_waves, "render_regen": best_regen,
            "render_grid": best_grid, "render_band": best_band, "render_pool": best_pool,
            "render_pool_ms": {str(k): round(v, 3) for k, v in res_pool.items()},
            "wf_stack": st, "wf_grid": g, "wf_trace_block": b,
            "wf_persistent": pm, "wf_refill": rf, "wf_trace_grid": tg or g, "wf_pool": pool,
            "wf_finish": fin,
            "megakernel_ms": {f"stack={k[0]},waves={k[1]},regen={k[2]},grid={k[3]},band={k[4]}":
                              round(v, 3) for k, v in res_mk.items()},
            "wavefront_ms": {str(k): round(v, 3) for k, v in res_wf.items()}}


def tune_batch(batch, cameras, spp, cfg: Configs, device: int = 0, tile_ws=(8, 16, 32, 64),
               grids=(512, 1024, 2048), chunks=(1, 8), stacks=(1, 2), waves=(4,),
               reps: int = 2, first_sample: int = 1 << 20, slots=(None,),
               slot_spp=None) -> dict:
    """Tune the batched renderer's tile shape, persistent grid, work ranges,
    stack placement and register budget, then, for the best of those, the
    sample indices per launch (slots; None keeps the current value), rendering
    slot_spp samples per scene (default: spp).  Every candidate renders the
    whole batch; all films of a stage must come out bit-identical before the
    fastest choice is applied."""
    ctx = get_context(device)
    sensors = [c.getSensor() for c in cameras]
    ref, res = None, {}
    for tw, g, ch, st, wv in itertools.product(tile_ws, grids, chunks, stacks, waves):
        ctx.set_tuning(batch_tile_w=tw, batch_grid=g, batch_chunks=ch, batch_stack=st,
                       batch_graph=1, batch_waves=wv)
        ms = []
        for r in range(reps + 1):
            for s in sensors:
                s.clear()
            out = batch.render(cameras, spp, cfg, first_sample=first_sample, readback=(r == 0))
            if r == 0:
                imgs = b"".join(s.data.tobytes() + s.weight.tobytes() for s in sensors)
                if ref is None:
                    ref = imgs
                elif imgs != ref:
                    raise AssertionError(f"batch tuning {(tw, g, ch, st, wv)} changed an image")
            else:
                ms.append(out["kernel_ms"])
        res[(tw, g, ch, st, wv)] = _median(ms)
    best = min(res, key=res.get)
    tw, g, ch, st, wv = best
    ctx.set_tuning(batch_tile_w=tw, batch_grid=g, batch_chunks=ch, batch_stack=st, batch_graph=1,
                   batch_waves=wv)
    sl0 = ctx.get_tuning()["batch_slots"]
    res_sl, ref2 = {}, None
    for sl in [sl0 if x is None else x for x in slots]:
        ctx.set_tuning(batch_slots=sl)
        ms = []
        for r in range(reps + 1):
            for s in sensors:
                s.clear()
            out = batch.render(cameras, slot_spp or spp, cfg, first_sample=first_sample,
                               readback=(r == 0))
            if r == 0:
                imgs = b"".join(s.data.tobytes() + s.weight.tobytes() for s in sensors)
                if ref2 is None:
                    ref2 = imgs
                elif imgs != ref2:
                    raise AssertionError(f"batch slots {sl} changed an image")
            else:
                ms.append(out["kernel_ms"])
        res_sl[sl] = _median(ms)
    best_sl = min(res_sl, key=res_sl.get)
    ctx.set_tuning(batch_slots=best_sl)
    return {"batch_tile_w": tw, "batch_grid": g, "batch_chunks": ch, "batch_stack": st,
            "batch_waves": wv, "batch_slots": best_sl,
            "slots_ms": {str(k): round(v, 3) for k, v in res_sl.items()},
            "timings_ms": {f"tile_w={k[0]},grid={k[1]},chunks={k[2]},stack={k[3]},waves={k[4]}":
                           round(v, 4) for k, v in res.items()}}


def tune_guided(integrator, scene: Scene, spp: int = 4, device: int = 0, stacks=(2, 3),
                waves=(4, 6, 8), grids=(2048, 4096), pools=(None,),
                first_sample: int = 1 << 20) -> dict:
    """Tune the fused guided megakernel (stack placement, register budget,
    persistent grid, paths per launch) on the scene's current camera with
    the integrator's guide as it stands (no training, so every candidate
    must render the same image bit for bit); applies the fastest."""
    ctx = get_context(device)
    sensor = scene.getCamera().getSensor()
    pool0 = ctx.get_tuning()["guided_pool"]
    res, ref = {}, None
    for st, wv, g, p in itertools.product(stacks, waves, grids,
                                          [pool0 if p is None else p for p in pools]):
        ctx.set_tuning(guided_fused=1, guided_stack=st, guided_waves=wv, guided_grid=g,
                       guided_pool=p)
        sensor.clear()
        integrator.dispatch(scene, spp, first_sample=first_sample, train=False, readback=False)
        sensor.clear()
        ms = integrator.dispatch(scene, spp, first_sample=first_sample, train=False)["kernel_ms"]
        img = sensor.data.tobytes() + sensor.weight.tobytes()
        if ref is None:
            ref = img
        elif img != ref:
            raise AssertionError(f"guided tuning {(st, wv, g, p)} changed the image")
        res[(st, wv, g, p)] = ms / spp
    best = min(res, key=res.get)
    st, wv, g, p = best
    ctx.set_tuning(guided_fused=1, guided_stack=st, guided_waves=wv, guided_grid=g, guided_pool=p)
    return {"guided_stack": st, "guided_waves": wv, "guided_grid": g, "guided_pool": p,
            "ms_per_spp": {str(k): round(v, 4) for k, v in res.items()}}


def tune_batch_wavefront(batch, cameras, spp, cfg: Configs, device: int = 0,
                         pools=(1 << 22, 1 << 23, 1 << 24), finishes=(1, 2, 65),
                         stacks=(1, 3), blocks=(128, 256), grids=(8192, 16384), reps: int = 1,
                         first_sample: int = 1 << 20) -> dict:
    """Tune the batched wavefront: paths in flight per pass (wf_pool, the
    queue capacity: sample indices of every pixel of the batch per pass),
    the bounce after which paths finish in one per-path launch, the stack
    placement, the traversal block size and the grids.  Every candidate
    renders the whole batch (graph replay), all films must come out
    bit-identical, and the fastest is applied."""
    ctx = get_context(device)
    sensors = [c.getSensor() for c in cameras]
    wcfg = Configs(**{**cfg.__dict__, "mode": L.MODE_WAVEFRONT})
    ref, res = None, {}
    for pool, fin, st, tb, g in itertools.product(pools, finishes, stacks, blocks, grids):
        ctx.set_tuning(wf_pool=pool, wf_finish=fin, wf_stack=st, wf_trace_block=tb,
                       wf_trace_grid=g, wf_shade_grid=g, batch_graph=1)
        ms = []
        for r in range(reps + 1):
            for s in sensors:
                s.clear()
            out = batch.render(cameras, spp, wcfg, first_sample=first_sample, readback=(r == 0))
            if r == 0:
                imgs = b"".join(s.data.tobytes() + s.weight.tobytes() for s in sensors)
                if ref is None:
                    ref = imgs
                elif imgs != ref:
                    raise AssertionError(f"batch wavefront tuning {(pool, fin, st, tb, g)} "
                                         "changed an image")
            else:
                ms.append(out["kernel_ms"])
        res[(pool, fin, st, tb, g)] = _median(ms)
    best = min(res, key=res.get)
    pool, fin, st, tb, g = best
    ctx.set_tuning(wf_pool=pool, wf_finish=fin, wf_stack=st, wf_trace_block=tb, wf_trace_grid=g,
                   wf_shade_grid=g, batch_graph=1)
    return {"wf_pool": pool, "wf_finish": fin, "wf_stack": st, "wf_trace_block": tb,
            "wf_grid": g, "ms": res[best],
            "timings_ms": {f"pool={k[0]},finish={k[1]},stack={k[2]},block={k[3]},grid={k[4]}":
                           round(v, 3) for k, v in res.items()}}
