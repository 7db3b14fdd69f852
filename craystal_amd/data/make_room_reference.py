"""Regenerates the occluded-light room's converged images: the room at
256 x 256 (room256_plain65536.npy, float32 RGB) or 1024 x 1024
(room1024_plain65536.npy, float16 RGB), max depth 16, Russian roulette from
depth 3, MIS, rendered by the plain megakernel with 65,536 samples per pixel
(sample indices 2^24 .. 2^24 + 65535).  The guided-vs-plain measurements
compare against them.  Needs a GPU: about 25 s at 256^2 and 6-7 minutes at
1024^2 on one MI355X with the path-regenerating megakernel.

Usage: python craystal_amd/data/make_room_reference.py [256|1024]"""
import sys
import time
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[2]))
import craystal_amd as cr  # noqa: E402
from craystal_amd import scenes  # noqa: E402

HERE = Path(__file__).resolve().parent


def reference_path(size: int) -> Path:
    return HERE / (f"room{size}_plain65536.npy")


def main(size: int = 256, spp: int = 65536, chunk: int = 1024):
    data, camd = scenes.occluded_room()
    scene = cr.Scene(data)
    sensor = cr.Sensor((size, size))
    scene.setCamera(scenes.make_camera(camd, sensor))
    cfg = cr.Configs(maxDepth=16, rrThreshold=0.05, useMIS=True, rrStartDepth=3)
    # Path regeneration, private stack, 64-VGPR budget: the launch shape the
    # bench's tuner picks on this scene (it does not change the image).
    cr.get_context(0).set_tuning(render_regen=1, render_grid=2048, render_stack=2,
                                 render_waves=8)
    integ = cr.PathTraceIntegrator(cfg)
    t0 = time.perf_counter()
    for done in range(0, spp, chunk):
        integ.dispatch(scene, min(chunk, spp - done), first_sample=(1 << 24) + done,
                       readback=False)
        print(f"{size}^2: {done + chunk} of {spp} spp, {time.perf_counter() - t0:.0f} s",
              flush=True)
    sensor.readbackDeviceData()
    img = sensor.createImage().astype(np.float32)
    np.save(reference_path(size), img.astype(np.float16) if size > 256 else img)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 256)
