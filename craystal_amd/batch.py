"""Batched rendering of many scenes (crh_batch_* in include/crayhip.h).

The scenes' device data are concatenated into one pool; every render call
runs, per sample index, one persistent megakernel launch over the tiles of
all scenes that still take samples and one film update over all of them,
replayed from a captured hipGraph.  Each scene's film ends bit-identical to
rendering that scene alone (PathTraceIntegrator.dispatch with the same
arguments).  The reference renders one scene per dispatch
(Core/Integrator.cu:233-262); this has no counterpart there.
"""
from __future__ import annotations

import ctypes as C

from . import _lib as L
from .core import Camera, Configs, Scene, get_context


class Batch:
    def __init__(self, scenes: list[Scene], device: int = 0):
        if not scenes:
            raise ValueError("a batch needs at least one scene")
        self.ctx = get_context(device)
        arr = (C.c_void_p * len(scenes))(*[s.handle.value for s in scenes])
        h = C.c_void_p()
        L.call("crh_batch_create", self.ctx.handle, arr, len(scenes), C.byref(h))
        self.handle = h
        self.n = len(scenes)

    def info(self) -> dict:
        i = L.crh_batch_info()
        L.call("crh_batch_get_info", self.handle, C.byref(i))
        return {k: getattr(i, k) for k, _ in L.crh_batch_info._fields_}

    def render(self, cameras: list[Camera], spp: list[int], configs: Configs | None = None,
               first_sample: int = 0, readback: bool = True) -> dict:
        """cameras[i] (with its sensor) views scene i; spp[i] samples each."""
        if len(cameras) != self.n or len(spp) != self.n:
            raise ValueError("one camera and one spp per scene")
        cfg = (configs or Configs()).to_c()
        cams = (L.crh_camera * self.n)(*[c.data for c in cameras])
        sensors = [c.getSensor() for c in cameras]
        if any(s is None for s in sensors):
            raise ValueError("every camera needs a sensor")
        films = (C.c_void_p * self.n)(*[s.handle.value for s in sensors])
        counts = (C.c_uint32 * self.n)(*[int(k) for k in spp])
        st = L.crh_stats()
        L.call("crh_batch_render", self.handle, cams, films, counts, C.byref(cfg), first_sample,
               C.byref(st))
        for s, k in zip(sensors, spp):
            s.setSPP(int(k))
            if readback:
                s.readbackDeviceData()
        return {"kernel_ms": st.kernel_ms, "total_ms": st.total_ms, "rays": st.rays,
                "launches": st.launches}

    def __del__(self):
        if getattr(self, "handle", None) and L._lib is not None:
            L._lib.crh_batch_destroy(self.handle)
            self.handle = None
