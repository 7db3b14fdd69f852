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
    t, spp, curve = 0.0, 0, []
    while t < budget_s and spp < max_spp:
        k, dt = render_chunk(spp)
        t += dt
        spp += k
        sensor.readbackDeviceData()
        e = relmse(sensor.createImage(), ref)
        curve.append((t, spp, e))
        if target > 0 and e <= target:
            break
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


def guided_vs_plain(data, camd, size, cfg: Configs, ref: np.ndarray, target: float = 0.01,
                    budget_s: float = 60.0, chunk: int = 8, train_fraction: float = 0.25,
                    train_batch: int = 16384, train_steps: int = 4, lr: float = 1e-2,
                    alpha: float = 0.5, max_spp: int = 1 << 20, seed: int = 7) -> dict:
    """Plain megakernel and guided wavefront, each for budget_s seconds of
    render time (the guided time includes its training steps, which run
    during the first train_fraction of its budget); relMSE against ref after
    every chunk of passes."""
    scene = Scene(data)
    sensor = Sensor(size)
    cam = scenes.make_camera(camd, sensor)
    scene.setCamera(cam)
    plain = PathTraceIntegrator(Configs(**{**cfg.__dict__, "mode": L.MODE_MEGAKERNEL}))
    plain.dispatch(scene, 1, first_sample=1 << 22, readback=False)     # warm up
    sensor.clear()

    def plain_chunk(first):
        st = plain.dispatch(scene, chunk, first_sample=first, readback=False)
        return chunk, st["kernel_ms"] / 1e3

    pcurve = _progressive(plain_chunk, sensor, ref, budget_s, 0.0, max_spp)

    P = data.concatenated()[0]
    guide = Guide(P.min(0), P.max(0), seed=seed)
    gcfg = Configs(**{**cfg.__dict__, "mode": L.MODE_WAVEFRONT})
    gi = GuidedIntegrator(gcfg, guide, alpha=alpha, lr=lr, train_batch=train_batch,
                          train_steps=train_steps)
    sensor.clear()
    gi.dispatch(scene, 1, first_sample=1 << 22, readback=False)       # warm up (no training)
    sensor.clear()
    train_until = train_fraction * budget_s
    state = {"t": 0.0, "train_s": 0.0}

    def guided_chunk(first):
        train = state["t"] < train_until
        st = gi.dispatch(scene, chunk, first_sample=first, train=train, readback=False)
        dt = st["kernel_ms"] / 1e3
        state["t"] += dt
        if train:
            state["train_s"] += dt
        return chunk, dt

    gcurve = _progressive(guided_chunk, sensor, ref, budget_s, 0.0, max_spp)
    t_plain, spp_plain = time_to_target(pcurve, target)
    t_guided, spp_guided = time_to_target(gcurve, target)
    x_plain, x_guided = extrapolate_time(pcurve, target), extrapolate_time(gcurve, target)
    px = size[0] * size[1]
    out = {
        "target_relmse": target, "budget_s": budget_s, "pixels": px,
        "plain": {"seconds_to_target": t_plain, "spp_to_target": spp_plain,
                  "extrapolated_seconds": x_plain, "final_relmse": pcurve[-1][2],
                  "final_spp": pcurve[-1][1],
                  "msamples_per_s": pcurve[-1][1] * px / pcurve[-1][0] / 1e6},
        "guided": {"seconds_to_target": t_guided, "spp_to_target": spp_guided,
                   "extrapolated_seconds": x_guided, "final_relmse": gcurve[-1][2],
                   "final_spp": gcurve[-1][1], "train_seconds": round(state["train_s"], 3),
                   "train_steps": guide.info()["steps"],
                   "msamples_per_s": gcurve[-1][1] * px / gcurve[-1][0] / 1e6},
        "curve_plain": [(round(t, 4), k, round(e, 5)) for t, k, e in pcurve[::max(1, len(pcurve) // 24)]],
        "curve_guided": [(round(t, 4), k, round(e, 5)) for t, k, e in gcurve[::max(1, len(gcurve) // 24)]],
    }
    tg = gcurve[-1][0]
    pe = [e for t, k, e in pcurve if t <= tg]
    out["equal_time"] = {"seconds": round(tg, 4), "plain_relmse": pe[-1] if pe else None,
                         "guided_relmse": gcurve[-1][2]}
    if t_plain and t_guided:
        out["speedup_measured"] = t_plain / t_guided
    if x_plain and x_guided:
        out["speedup_extrapolated"] = x_plain / x_guided
    return out
