"""The C++ host API (include/craystal/CRaystal.h) compiles against the
C-ABI and behaves like the reference's classes (tests/cpp/api_test.cpp)."""
import subprocess
from pathlib import Path

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
