"""Command line renderer with the reference CLI's options (CRayli.cpp:16-42):

    python -m craystal_amd.cli scene.xml [--spp 16] [--no-mis] [-o craystal.exr]

Imports the XML + OBJ scene, renders `spp` samples per pixel with the
integrator configuration the reference's kernel runs with (max depth 10,
Russian roulette threshold 0.05 after depth 5; MIS unless --no-mis) and
writes the image as a half-float EXR (B, G, R channels).
"""
from __future__ import annotations

import argparse
import sys

from . import _lib as L
from .core import Configs, PathTraceIntegrator
from .imageio import write_exr
from .importer import Importer


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="CRaystal renderer (MI355X)")
    ap.add_argument("modelXML", help="Path to model XML file")
    ap.add_argument("--spp", type=int, default=16, help="Samples per pixel")
    ap.add_argument("--no-mis", action="store_true",
                    help="Disable multiple importance sampling")
    ap.add_argument("-o", "--output", default="craystal.exr", help="output EXR path")
    ap.add_argument("--wavefront", action="store_true", help="use the wavefront integrator")
    a = ap.parse_args(argv)
    L.load()
    scene = Importer().load(a.modelXML)
    cfg = Configs.reference(useMIS=not a.no_mis)
    if a.wavefront:
        cfg.mode = L.MODE_WAVEFRONT
    st = PathTraceIntegrator(cfg).dispatch(scene, a.spp)
    write_exr(a.output, scene.getCamera().getSensor().createImage())
    print(f"wrote {a.output}: {a.spp} spp, {st['kernel_ms']:.1f} ms, {st['rays']} rays")
    return 0


if __name__ == "__main__":
    sys.exit(main())
