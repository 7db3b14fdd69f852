"""Time-to-relMSE of the guided wavefront path vs the plain megakernel.

relMSE(I, R) = mean over pixels and channels of (I - R)^2 / (R^2 + 1e-2).
The reference R is a long plain render of the same scene.  Both methods
render progressively (sample indices 0, 1, 2, ...) into a film; after every
chunk of passes the film is read back and compared with R.  Rendering time
is the GPU time of the render calls (the readback and the comparison are
outside it).  The guided method trains its network during the first
`train_fraction` of the time budget and keeps rendering (without training)
afterwards; every guided sample, trained or not, is unbiased and counts.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib as L
from .core import Configs, PathTraceIntegrator, Scene, Sensor
from .guiding import Guide, GuidedIntegrator
from . import scenes

EPS = 1e-2


def relmse(img: np.ndarray, ref: np.ndarray) -> float:
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + EPS)))


def relmse_trimmed(img: np.ndarray, ref: np.ndarray, frac: float = 1e-3) -> float:
    """relMSE without the `frac` of pixels with the largest error: a
    diagnostic of the bulk of the image that single extreme samples do not
    move (the metric itself is relmse)."""
    e = ((img - ref) ** 2 / (ref ** 2 + EPS)).mean(axis=-1).ravel()
    keep = len(e) - int(len(e) * frac)
    return float(np.sort(e)[:keep].mean()) if keep > 0 else float("nan")


class Curve(list):
    """[(t, spp, relmse)] with the final image's trimmed relMSE."""
    trimmed = None


def reference_image(scene: Scene, camd: dict, size, cfg: Configs, spp: int,
                    chunk: int = 256, first: int = 1 << 24) -> np.ndarray:
    sensor = Sensor(size)
    cam = scenes.make_camera(camd, sensor)
    scene.setCamera(cam)
    integ = PathTraceIntegrator(cfg)
    done = 0
    while done < spp:
        k = min(chunk, spp - done)
        integ.dispatch(scene, k, first_sample=first + done, readback=False)
        done += k
    sensor.readbackDeviceData()
    return sensor.createImage()


def _progressive(render_chunk, sensor: Sensor, ref: np.ndarray, budget_s: float, target: float,
                 max_spp: int):
    """Runs render_chunk(first_sample) -> (spp rendered, seconds) until the
    relMSE target (0 = never) or the time budget; returns the curve
    [(t, spp, relmse)]."""
    t, spp, curve = 0.0, 0, Curve()
    img = None
    while t < budget_s and spp < max_spp:
        k, dt = render_chunk(spp)
        t += dt
        spp += k
        sensor.readbackDeviceData()
        img = sensor.createImage()
        e = relmse(img, ref)
        curve.append((t, spp, e))
        if target > 0 and e <= target:
            break
    if img is not None:
        curve.trimmed = relmse_trimmed(img, ref)
    return curve


def time_to_target(curve, target):
    for t, spp, e in curve:
        if e <= target:
            return t, spp
    return None, None


def extrapolate_time(curve, target, tail: float = 0.5):
    """Time at which relMSE reaches target under relMSE(t) = a / t, with a
    fitted (least squares in a) on the last `tail` fraction of the curve."""
    pts = [(t, e) for t, _, e in curve[int(len(curve) * (1 - tail)):] if t > 0 and e > 0]
    if not pts:
        return None
    inv = np.array([1.0 / t for t, _ in pts])
    err = np.array([e for _, e in pts])
    a = float(inv @ err / (inv @ inv))
    return a / target


def plain_curve(scene: Scene, sensor: Sensor, cfg: Configs, ref: np.ndarray, budget_s: float,
                chunk: int, mode: int = L.MODE_MEGAKERNEL, max_spp: int = 1 << 20,
                first: int = 0):
    """Progressive plain render (megakernel or wavefront) for budget_s
    seconds of render time; relMSE against ref after every chunk."""
    plain = PathTraceIntegrator(Configs(**{**cfg.__dict__, "mode": mode}))
    plain.dispatch(scene, 1, first_sample=1 << 22, readback=False)     # warm up
    sensor.clear()

    def run(done):
        st = plain.dispatch(scene, chunk, first_sample=first + done, readback=False)
        return chunk, st["kernel_ms"] / 1e3

    return _progressive(run, sensor, ref, budget_s, 0.0, max_spp)


def guided_curve(scene: Scene, sensor: Sensor, data, cfg: Configs, ref: np.ndarray,
                 budget_s: float, chunk: int, train_fraction: float = 0.25,
                 train_batch: int = 4096, train_steps: int = 4, lr: float = 1e-2,
                 alpha: float = 0.5, max_depth: int = 0, materials: int = 0, target: int = 0,
                 seed: int = 7, max_spp: int = 1 << 20, first: int = 0):
    """Progressive guided wavefront render for budget_s seconds of render
    time, the guide training during the first train_fraction of it (its
    training steps count as render time).  Returns (curve, info)."""
    P = data.concatenated()[0]
    guide = Guide(P.min(0), P.max(0), seed=seed)
    gcfg = Configs(**{**cfg.__dict__, "mode": L.MODE_WAVEFRONT})
    gi = GuidedIntegrator(gcfg, guide, alpha=alpha, lr=lr, train_batch=train_batch,
                          train_steps=train_steps, max_depth=max_depth, materials=materials,
                          target=target)
    gi.dispatch(scene, 1, first_sample=1 << 22, readback=False)       # warm up (no training)
    sensor.clear()
    train_until = train_fraction * budget_s
    state = {"t": 0.0, "train_s": 0.0}

    def run(done):
        train = state["t"] < train_until
        st = gi.dispatch(scene, chunk, first_sample=first + done, train=train, readback=False)
        dt = st["kernel_ms"] / 1e3
        state["t"] += dt
        if train:
            state["train_s"] += dt
        return chunk, dt

    curve = _progressive(run, sensor, ref, budget_s, 0.0, max_spp)
    return curve, {"train_seconds": round(state["train_s"], 3),
                   "train_steps": guide.info()["steps"]}


def summarize(curve, target: float, pixels: int) -> dict:
    t_hit, spp_hit = time_to_target(curve, target)
    return {"seconds_to_target": t_hit, "spp_to_target": spp_hit,
            "extrapolated_seconds": extrapolate_time(curve, target),
            "final_relmse": curve[-1][2], "final_spp": curve[-1][1],
            "final_relmse_trimmed": getattr(curve, "trimmed", None),
            "final_seconds": curve[-1][0],
            "msamples_per_s": curve[-1][1] * pixels / curve[-1][0] / 1e6,
            "curve": [(round(t, 4), k, round(e, 5)) for t, k, e in curve[::max(1, len(curve) // 24)]]}


def guided_vs_plain(data, camd, size, cfg: Configs, ref: np.ndarray, target: float = 0.01,
                    budget_s: float = 60.0, chunk: int = 8, plain_wavefront: bool = True,
                    **guided_kw) -> dict:
    """Plain megakernel, plain wavefront and guided wavefront, each for
    budget_s seconds of render time; relMSE against ref after every chunk."""
    scene = Scene(data)
    sensor = Sensor(size)
    scene.setCamera(scenes.make_camera(camd, sensor))
    px = size[0] * size[1]
    out = {"target_relmse": target, "budget_s": budget_s, "pixels": px}
    pc = plain_curve(scene, sensor, cfg, ref, budget_s, chunk)
    out["plain"] = summarize(pc, target, px)
    if plain_wavefront:
        wc = plain_curve(scene, sensor, cfg, ref, budget_s, chunk, mode=L.MODE_WAVEFRONT)
        out["plain_wavefront"] = summarize(wc, target, px)
    gc, info = guided_curve(scene, sensor, data, cfg, ref, budget_s, chunk, **guided_kw)
    out["guided"] = {**summarize(gc, target, px), **info}
    p, g = out["plain"], out["guided"]
    if p["seconds_to_target"] and g["seconds_to_target"]:
        out["speedup_measured"] = p["seconds_to_target"] / g["seconds_to_target"]
    if p["extrapolated_seconds"] and g["extrapolated_seconds"]:
        out["speedup_extrapolated"] = p["extrapolated_seconds"] / g["extrapolated_seconds"]
    tg = gc[-1][0]
    pe = [e for t, k, e in pc if t <= tg]
    out["equal_time"] = {"seconds": round(tg, 4), "plain_relmse": pe[-1] if pe else None,
                         "guided_relmse": gc[-1][2]}
    return out
