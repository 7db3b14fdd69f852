"""Regenerates room256_plain65536.npy: the occluded-light room at 256 x 256,
max depth 16, Russian roulette from depth 3, MIS, rendered by the plain
megakernel with 65,536 samples per pixel (sample indices 2^24 .. 2^24 +
65535), as float32 RGB.  The guided-vs-plain measurements compare against
it.  Needs a GPU (about 200 s on one MI355X)."""
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[2]))
import craystal_amd as cr  # noqa: E402
from craystal_amd import scenes  # noqa: E402


def main(out=Path(__file__).resolve().parent / "room256_plain65536.npy", spp=65536):
    data, camd = scenes.occluded_room()
    scene = cr.Scene(data)
    sensor = cr.Sensor((256, 256))
    scene.setCamera(scenes.make_camera(camd, sensor))
    cfg = cr.Configs(maxDepth=16, rrThreshold=0.05, useMIS=True, rrStartDepth=3)
    integ = cr.PathTraceIntegrator(cfg)
    for done in range(0, spp, 1024):
        integ.dispatch(scene, min(1024, spp - done), first_sample=(1 << 24) + done,
                       readback=False)
    sensor.readbackDeviceData()
    np.save(out, sensor.createImage().astype(np.float32))


if __name__ == "__main__":
    main()
