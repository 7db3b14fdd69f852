"""The C++ host API (include/craystal/CRaystal.h) compiles against the
C-ABI and behaves like the reference's classes (tests/cpp/api_test.cpp)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "tests" / "cpp" / "_build" / "api_test"


def _exe():
    """Built by craystal_amd.build (as are the library's own objects)."""
    if not EXE.exists():
        from craystal_amd import build
        build.build_cpp_tests()
    return EXE


def _env():
    import os
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = str(ROOT / "craystal_amd") + ":" + env.get("LD_LIBRARY_PATH", "")
    return env


def test_cpp_host_api():
    r = subprocess.run([str(_exe())], capture_output=True, text=True, env=_env())
    assert r.returncode == 0, r.stderr + r.stdout
    assert "cpp host api ok" in r.stdout


@pytest.mark.gpu
def test_cpp_gpu_api(gpu):
    r = subprocess.run([str(_exe()), "--gpu"], capture_output=True, text=True, env=_env(),
                       timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    assert "cpp gpu api ok" in r.stdout


SHAPE = ROOT / "tests" / "cpp" / "_build" / "reference_shape_test"


def _shape_exe():
    if not SHAPE.exists():
        from craystal_amd import build
        build.build_cpp_tests()
    return SHAPE


def test_reference_shaped_host_code_compiles_and_runs():
    """tests/cpp/reference_shape_test.cpp is written like the reference's
    caller (CRayli.cpp:7-42: Float3 / UInt2 members, Spectrum::fromRGB,
    Image::Ref Sensor::createImage(), Image::writeEXR); it builds against
    include/craystal and its host part runs."""
    r = subprocess.run([str(_shape_exe())], capture_output=True, text=True, env=_env())
    assert r.returncode == 0, r.stderr + r.stdout
    assert "reference-shaped host code compiles" in r.stdout


@pytest.mark.gpu
def test_reference_shaped_host_code_renders(gpu, tmp_path):
    from craystal_amd.imageio import read_exr
    out = tmp_path / "craystal.exr"
    r = subprocess.run([str(_shape_exe()), str(out)], capture_output=True, text=True,
                       env=_env(), timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    img, hdr = read_exr(out)
    assert img.shape == (30, 40, 3) and np.isfinite(img).all() and img.mean() > 0.0
