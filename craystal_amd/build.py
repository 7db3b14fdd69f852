"""Build libcrayhip.so (gfx950 HIP kernels + C-ABI) in-tree.

Host C++ (api.cpp, bvh_build.cpp, spectral.cpp) is compiled with g++ against the HIP
runtime headers; the kernels with hipcc for gfx950.  Everything is built with
-ffp-contract=off so that traversal arithmetic is bit-identical to the
host-side restatement of the reference.  Run as ``python -m craystal_amd.build``.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "libcrayhip.so"
OBJ = PKG / "_build"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = os.environ.get("CRH_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
    "-ffp-contract=off",
    # Packed FP32 (v_pk_*_f32) generation off.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
HOST_FLAGS = [
    "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function", f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__",
]


def _run(cmd: list[str]) -> None:
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _run_all(cmds: list[list[str]], jobs: int | None = None) -> None:
    """The compile commands (independent translation units), at most `jobs`
    at a time (default: the CPU count, at most 8)."""
    jobs = jobs or max(1, min(8, os.cpu_count() or 1))
    running: list[tuple[subprocess.Popen, list[str]]] = []
    failed: list[tuple[int, list[str]]] = []

    def reap(block: bool) -> None:
        for p, c in list(running):
            rc = p.wait() if block else p.poll()
            if rc is not None:
                running.remove((p, c))
                if rc != 0:
                    failed.append((rc, c))
                if block:
                    return

    for c in cmds:
        while len(running) >= jobs:
            reap(block=False)
            if len(running) >= jobs:
                running[0][0].wait()
        print(" ".join(c), flush=True)
        running.append((subprocess.Popen(c), c))
    while running:
        reap(block=True)
    if failed:
        raise subprocess.CalledProcessError(failed[0][0], failed[0][1])


def build(verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    objs, cmds = [], []
    hipcc = str(ROCM / "bin" / "hipcc")

    def hip(src: str, obj: str, *defs: str) -> None:
        o = OBJ / obj
        cmds.append([hipcc, *HIPCC_FLAGS, *defs, "-c", str(CSRC / src), "-o", str(o)])
        objs.append(str(o))

    # Longest compiles first.
    # The fused guided kernel: one object per stack mode, then the rest.
    for part in range(4):
        hip("guided_mk.hip", f"guided_mk_{part}.hip.o", f"-DCRH_GMK_PART={part}")
    hip("wavefront.hip", "wavefront.hip.o")
    for sm in range(3):
        hip("wf_finish.hip", f"wf_finish_{sm}.hip.o", f"-DCRH_WF_SM={sm}")
    # The megakernel families, one object per (family, stack mode), so that
    # their many register-budget variants compile in parallel.
    for fam in range(3):
        for sm in range(3):
            hip("mk_inst.hip", f"mk_inst_{fam}_{sm}.hip.o", f"-DCRH_MK_FAMILY={fam}",
                f"-DCRH_MK_SM={sm}")
    for src in ("kernels.hip", "guiding.hip", "noalloc.hip", "bvh_device.hip"):
        hip(src, src + ".o")
    for src in ("api.cpp", "bvh_build.cpp", "spectral.cpp"):
        o = OBJ / (src + ".o")
        cmds.append(["g++", *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(o)])
        objs.append(str(o))
    _run_all(cmds)
    tmp = OUT.with_suffix(".so.tmp")
    _run([str(ROCM / "bin" / "hipcc"), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *objs, "-lpthread"])
    os.replace(tmp, OUT)
    return OUT


def build_cpp_tests() -> Path:
    """Compile tests/cpp/api_test.cpp and import_test.cpp against include/ and
    libcrayhip.so."""
    root = PKG.parent
    outdir = root / "tests" / "cpp" / "_build"
    outdir.mkdir(parents=True, exist_ok=True)
    for name in ("api_test", "import_test", "reference_shape_test"):
        _run(["g++", "-O1", "-std=c++17", f"-I{root / 'include'}",
              str(root / "tests" / "cpp" / f"{name}.cpp"), f"-L{PKG}", "-lcrayhip",
              f"-Wl,-rpath,{PKG}", "-lz", "-o", str(outdir / name)])
    return outdir / "api_test"


def build_cli() -> Path:
    """The crayli command line renderer next to libcrayhip.so."""
    out = PKG / "crayli"
    _run(["g++", "-O2", "-std=c++17", f"-I{PKG.parent / 'include'}", str(CSRC / "crayli.cpp"),
          f"-L{PKG}", "-lcrayhip", "-Wl,-rpath,$ORIGIN", "-lz", "-o", str(out)])
    return out


if __name__ == "__main__":
    build()
    build_cli()
    build_cpp_tests()
    sys.exit(0)
