"""The fused guided megakernel (guide inference on MFMA inside the persistent
path kernel, crh_tuning.guided_fused = 1) against the guided wavefront
(guided_fused = 2), whose estimator the oracle pins
(test_gpu_guiding.test_guided_film_matches_oracle_with_a_frozen_guide).

For the same guide both make every draw and every float operation of a
path in the same order, so the films are compared bit for bit: a frozen
guide (the output layer's bias only) and a randomly initialised one, with
and without MIS, guiding all vertices or a material / depth subset, and
every stack / register-budget variant of the fused kernel.  Training in
the fused path is checked by its steps, losses and the variance it
removes; unbiasedness by image means against the plain megakernel."""
import numpy as np
import pytest

import craystal_amd as cr
from craystal_amd import _lib as L
from craystal_amd import scenes
from craystal_amd.guiding import Guide, GuidedIntegrator

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _film(sensor):
    return sensor.data.tobytes(), sensor.weight.tobytes()


def _frozen_theta(seed=4):
    rng = np.random.default_rng(seed)
    o = rng.normal(size=L.GUIDE_OUTPUTS) * 1.5
    for j in range(8):
        o[5 * j + 4] = rng.uniform(-1.0, 4.0)
    theta = np.zeros(L.GUIDE_PARAMS, np.float32)
    theta[-L.GUIDE_OUTPUTS:] = o
    return theta


@pytest.fixture
def tuning_restored():
    ctx = cr.get_context(0)
    saved = ctx.get_tuning()
    yield ctx
    ctx.set_tuning(**saved)


def _render(ctx, scene, sensor, gi, fused, spp, first, **tun):
    ctx.set_tuning(guided_fused=1 if fused else 2, **tun)
    sensor.clear()
    st = gi.dispatch(scene, spp, first_sample=first, train=False)
    return _film(sensor), st


def _cornell_principled():
    data, camd = scenes.cornell_box()
    data.materials[0] = scenes.principled((0.7, 0.6, 0.5), 0.35, 0.6)
    return data, camd


@pytest.mark.parametrize("mis", [True, False])
@pytest.mark.parametrize("guide_kind", ["frozen", "random"])
def test_fused_film_equals_wavefront(gpu, tuning_restored, mis, guide_kind):
    ctx = tuning_restored
    data, camd = _cornell_principled()
    P = np.array(data.concatenated()[0])
    guide = Guide(P.min(0), P.max(0), seed=9)
    if guide_kind == "frozen":
        guide.set_params(_frozen_theta())
    scene = cr.Scene(data)
    sensor = cr.Sensor((48, 40))
    scene.setCamera(scenes.make_camera(camd, sensor))
    cfg = cr.Configs(maxDepth=8, rrThreshold=0.05, useMIS=mis, rrStartDepth=3)
    gi = GuidedIntegrator(cfg, guide, alpha=0.5)
    a, sa = _render(ctx, scene, sensor, gi, True, 8, 5)
    b, _ = _render(ctx, scene, sensor, gi, False, 8, 5)
    _say(f"{guide_kind} guide, MIS {mis}: fused {sa['kernel_ms']:.2f} ms")
    assert a[1] == b[1], "sample weights differ"
    assert a[0] == b[0], "fused and wavefront guided films differ"
    img = sensor.createImage()
    assert np.isfinite(img).all() and img.mean() > 0


@pytest.mark.parametrize("materials,max_depth", [(1, 0), (2, 0), (3, 2), (1, 1)])
def test_fused_film_equals_wavefront_on_subsets(gpu, tuning_restored, materials, max_depth):
    """Guiding only Lambertian / only Principled vertices, or only the first
    vertices of each path, on the occluded room (Lambertian, Principled and
    Phong surfaces), at a size that is not a multiple of the 16 x 16 tile."""
    ctx = tuning_restored
    data, camd = scenes.occluded_room(n_objects=12, subdiv=1)
    P = np.array(data.concatenated()[0])
    guide = Guide(P.min(0), P.max(0), seed=21)
    scene = cr.Scene(data)
    sensor = cr.Sensor((70, 37))
    scene.setCamera(scenes.make_camera(camd, sensor))
    cfg = cr.Configs(maxDepth=12, rrThreshold=0.05, useMIS=True, rrStartDepth=3)
    gi = GuidedIntegrator(cfg, guide, alpha=0.35, materials=materials, max_depth=max_depth)
    a, _ = _render(ctx, scene, sensor, gi, True, 6, 77)
    b, _ = _render(ctx, scene, sensor, gi, False, 6, 77)
    assert a == b, "fused and wavefront guided films differ"


def test_fused_launch_variants_give_one_film(gpu, tuning_restored):
    """Stack placement, register budget, grid and pool size change the launch
    only: the film stays bit-identical."""
    ctx = tuning_restored
    data, camd = scenes.occluded_room(n_objects=12, subdiv=1)
    P = np.array(data.concatenated()[0])
    guide = Guide(P.min(0), P.max(0), seed=2)
    scene = cr.Scene(data)
    sensor = cr.Sensor((64, 64))
    scene.setCamera(scenes.make_camera(camd, sensor))
    cfg = cr.Configs(maxDepth=10, rrThreshold=0.05, useMIS=True, rrStartDepth=3)
    gi = GuidedIntegrator(cfg, guide, alpha=0.5, materials=1)
    ref, _ = _render(ctx, scene, sensor, gi, True, 12, 300, guided_stack=1, guided_waves=4,
                     guided_grid=2048, guided_pool=1 << 23)
    for stack in (1, 2, 3):
        for waves in (1, 4, 8):
            for grid, pool in ((2048, 1 << 23), (97, 5000)):
                got, st = _render(ctx, scene, sensor, gi, True, 12, 300, guided_stack=stack,
                                  guided_waves=waves, guided_grid=grid, guided_pool=pool)
                _say(f"stack {stack} waves {waves} grid {grid} pool {pool}: "
                     f"{st['kernel_ms']:.2f} ms")
                assert got == ref, (stack, waves, grid, pool)


@pytest.mark.parametrize("mis", [True, False])
def test_fused_film_equals_wavefront_width128(gpu, tuning_restored, mis):
    """The 128-wide network (hidden layers of 4 x 8 MFMA operand blocks,
    weights read per 16-column block) in the fused kernel and in the guided
    wavefront's inference kernel: bit-identical films for a randomly
    initialised guide, at every stack placement and register budget tried."""
    ctx = tuning_restored
    data, camd = scenes.occluded_room(n_objects=12, subdiv=1)
    P = np.array(data.concatenated()[0])
    guide = Guide(P.min(0), P.max(0), seed=13, width=128)
    assert guide.info()["width"] == 128 and len(guide.get_params()) == L.GUIDE_PARAMS_128
    scene = cr.Scene(data)
    sensor = cr.Sensor((56, 48))
    scene.setCamera(scenes.make_camera(camd, sensor))
    cfg = cr.Configs(maxDepth=10, rrThreshold=0.05, useMIS=mis, rrStartDepth=3)
    gi = GuidedIntegrator(cfg, guide, alpha=0.5)
    ref, _ = _render(ctx, scene, sensor, gi, False, 6, 41)
    for stack, waves in ((2, 6), (3, 8), (1, 4)):
        got, st = _render(ctx, scene, sensor, gi, True, 6, 41, guided_stack=stack,
                          guided_waves=waves)
        _say(f"width 128, stack {stack} waves {waves}: {st['kernel_ms']:.2f} ms")
        assert got == ref, (stack, waves)
    img = sensor.createImage()
    assert np.isfinite(img).all() and img.mean() > 0


def test_fused_training_runs_and_is_unbiased(gpu, tuning_restored):
    """Training from the fused kernel's records: steps run with finite
    losses, and trained / untrained guided image means agree with the plain
    megakernel's within 3 sigma."""
    ctx = tuning_restored
    ctx.set_tuning(guided_fused=1)
    data, camd = scenes.occluded_room(n_objects=12, subdiv=1)
    scene = cr.Scene(data)
    sensor = cr.Sensor((64, 64))
    scene.setCamera(scenes.make_camera(camd, sensor))
    cfg = cr.Configs(maxDepth=8, rrThreshold=0.05, useMIS=True, rrStartDepth=3)
    P = np.array(data.concatenated()[0])
    guide = Guide(P.min(0), P.max(0), seed=3)
    gi = GuidedIntegrator(cfg, guide, train_batch=256, train_steps=2)
    for k in range(6):
        sensor.clear()
        st = gi.dispatch(scene, 1, first_sample=10000 + k, train=True)
    assert st["train_samples"] > 0
    info = guide.info()
    assert 1 <= info["steps"] <= 12 and info["last_samples"] == st["train_samples"]
    losses = guide.losses()
    assert len(losses) == info["steps"] and np.isfinite(losses).all()
    plain, guided = [], []
    for k in range(24):
        sensor.clear()
        cr.PathTraceIntegrator(cfg).dispatch(scene, 1, first_sample=k)
        plain.append(sensor.createImage().mean())
        sensor.clear()
        gi.dispatch(scene, 1, first_sample=500 + k)
        guided.append(sensor.createImage().mean())
    plain, guided = np.array(plain), np.array(guided)
    assert np.isfinite(guided).all()
    se = np.sqrt(plain.var(ddof=1) / len(plain) + guided.var(ddof=1) / len(guided))
    assert abs(plain.mean() - guided.mean()) <= 3 * se, (plain.mean(), guided.mean(), se)


def test_fused_trained_guide_lowers_variance(gpu, tuning_restored):
    """The lampshade room (light only through the ceiling patch around an
    upward lamp): after training in the fused kernel, the guided per-pixel
    variance at equal spp is clearly below the plain megakernel's."""
    ctx = tuning_restored
    ctx.set_tuning(guided_fused=1)
    data, camd = scenes.lampshade_room()
    scene = cr.Scene(data)
    sensor = cr.Sensor((64, 64))
    scene.setCamera(scenes.make_camera(camd, sensor))
    cfg = cr.Configs(maxDepth=8, rrThreshold=0.05, useMIS=True, rrStartDepth=3)
    P = np.array(data.concatenated()[0])
    spp = 4

    def img(integ, k):
        sensor.clear()
        integ.dispatch(scene, spp, first_sample=1000 + spp * k)
        return sensor.createImage().copy()

    def relvar(render, n=24):
        imgs = np.array([render(k) for k in range(n)])
        m = imgs.mean(0)
        return float((imgs.var(0, ddof=1) / (m * m + 1e-2)).mean()), imgs.mean(axis=(1, 2, 3))

    v_plain, m_plain = relvar(lambda k: img(cr.PathTraceIntegrator(cfg), k))
    guide = Guide(P.min(0), P.max(0), seed=5)
    gi = GuidedIntegrator(cfg, guide, alpha=0.5, lr=1e-2, train_batch=4096, train_steps=4)
    for k in range(48):
        sensor.clear()
        gi.dispatch(scene, 8, first_sample=100000 + 8 * k, train=True)
    assert guide.info()["steps"] > 50
    v_guided, m_guided = relvar(lambda k: img(gi, k))
    _say(f"relative variance at {spp} spp: plain {v_plain:.4f}, fused guided {v_guided:.4f}")
    assert v_guided < 0.8 * v_plain, (v_guided, v_plain)
    se = np.sqrt(m_plain.var(ddof=1) / len(m_plain) + m_guided.var(ddof=1) / len(m_guided))
    assert abs(m_plain.mean() - m_guided.mean()) <= 3 * se
