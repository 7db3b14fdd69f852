"""Object lifetime across interpreter shutdown.

At exit CPython may finalise a context before the scenes, films and guides
made on it (module globals become cyclic garbage and their finalisers run in
list order).  Each native object holds a counted reference to its context,
so any order is safe; this runs a script that leaves everything alive at
exit and checks the interpreter exits cleanly.
"""
import subprocess
import sys
import textwrap
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]

SCRIPT = textwrap.dedent("""
    import sys
    sys.path.insert(0, {root!r})
    import numpy as np
    import craystal_amd as cr
    from craystal_amd import _lib as L, scenes
    from craystal_amd.guiding import Guide, GuidedIntegrator
    data, camd = scenes.cornell_box(seed=1)
    scene = cr.Scene(data)
    sensor = cr.Sensor((32, 32))
    cam = scenes.make_camera(camd, sensor)
    scene.setCamera(cam)
    cfg = cr.Configs(maxDepth=4, mode=L.MODE_WAVEFRONT)
    P = np.array(data.concatenated()[0])
    guide = Guide(P.min(0), P.max(0), seed=1)
    GuidedIntegrator(cfg, guide, train_batch=256, train_steps=1).dispatch(scene, 1, train=True)
    cr.PathTraceIntegrator(cfg).dispatch(scene, 1)
    buf = cr.DeviceBuffer(1024)
    ctx = cr.get_context(0)
    print("done", flush=True)
""")


@pytest.mark.gpu
def test_exit_with_live_objects_is_clean(gpu):
    r = subprocess.run([sys.executable, "-c", SCRIPT.format(root=str(ROOT))],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    assert "done" in r.stdout
