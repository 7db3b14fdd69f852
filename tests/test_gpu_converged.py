"""Rendered images against a converged image, on the Cornell box (RGB, and
spectral with the CIE XYZ film), 64 x 64, max depth 8, MIS.

The GPU megakernel at 65,536 spp is the converged image.  The CPU oracle's
restatement of pathTraceKernel (Core/Integrator.cu:142-227) renders 1,024 spp
of other sample indices, in 16 chunks, so that the spread of its chunk means
gives the standard error of its image mean.  The bar:

* per-channel image means of the oracle within 3 sigma of the converged
  image's;
* relMSE (eps 1e-2) of the oracle's image against the converged image below
  0.02;
* that relMSE equal, within 0.8-1.25x, to the relMSE of a GPU image of the
  same spp with a third set of sample indices: the CPU path and the GPU path
  have the same error against the converged image, not only the same mean.

The scene's materials are Lambertian, so the estimator's weights are bounded
and both statistics are stable at these sample counts.
"""
import numpy as np
import pytest

import craystal_amd as cr
from craystal_amd import scenes

pytestmark = pytest.mark.gpu

SIZE = (64, 64)
CONVERGED_SPP = 65536
SPP = 1024
CHUNKS = 16


def _relmse(img, ref, eps=1e-2):
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + eps)))


@pytest.mark.parametrize("spectral", [False, True])
def test_cpu_path_against_converged_image(gpu, orc, spectral):
    data, camd = scenes.cornell_box_spectral() if spectral else scenes.cornell_box()
    cfg = cr.Configs(maxDepth=8, rrThreshold=0.05, useMIS=True, rrStartDepth=3,
                     spectral=spectral)
    scene = cr.Scene(data)
    sensor = cr.Sensor(SIZE)
    cam = scenes.make_camera(camd, sensor)
    scene.setCamera(cam)

    def gpu_image(spp, first):
        sensor.clear()
        cr.PathTraceIntegrator(cfg).dispatch(scene, spp, first_sample=first)
        return sensor.createImage().astype(np.float64)

    conv = gpu_image(CONVERGED_SPP, 1 << 20)
    other = gpu_image(SPP, 1 << 21)
    assert np.isfinite(conv).all() and conv[..., 1].mean() > 0.01

    osc = orc.Scene(data)
    per = SPP // CHUNKS
    s_tot, w_tot, means = 0.0, 0.0, []
    for k in range(CHUNKS):
        s, w, _ = osc.render(cam, SIZE, cfg, k * per, per)
        means.append((s / w[..., None]).reshape(-1, 3).mean(0))
        s_tot, w_tot = s_tot + s.astype(np.float64), w_tot + w.astype(np.float64)
    cpu = s_tot / w_tot[..., None]
    means = np.array(means)
    se = means.std(0, ddof=1) / np.sqrt(CHUNKS)
    dm = np.abs(cpu.reshape(-1, 3).mean(0) - conv.reshape(-1, 3).mean(0))
    assert (dm <= 3 * se + 1e-7).all(), (dm, se)

    e_cpu, e_gpu = _relmse(cpu, conv), _relmse(other, conv)
    assert e_cpu < 0.02, e_cpu
    assert 0.8 < e_cpu / e_gpu < 1.25, (e_cpu, e_gpu)
