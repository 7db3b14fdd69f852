"""Host wrapper of the learned path guide (crh_guide_* in include/crayhip.h).

The network maps (position, shading normal, outgoing direction) to an
8-lobe von Mises-Fisher mixture over incident directions; it runs and trains
on the GPU (BF16 MFMA, FP32 accumulation, Adam).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .core import get_context


class Guide:
    def __init__(self, bmin, bmax, seed: int = 1234, device: int = 0, width: int = 64):
        """width: the hidden width of the network, 64 or 128."""
        self.ctx = get_context(device)
        lo = np.ascontiguousarray(bmin, np.float32)
        hi = np.ascontiguousarray(bmax, np.float32)
        h = C.c_void_p()
        L.call("crh_guide_create_width", self.ctx.handle, L.ptr(lo), L.ptr(hi), seed, width,
               C.byref(h))
        self.handle = h
        self.bmin, self.bmax = lo, hi
        self.width = width
        self.n_params = L.GUIDE_PARAMS_128 if width == 128 else L.GUIDE_PARAMS

    def set_params(self, theta: np.ndarray):
        t = np.ascontiguousarray(theta, np.float32).reshape(-1)
        L.call("crh_guide_set_params", self.handle, L.ptr(t), len(t))

    def get_params(self) -> np.ndarray:
        t = np.zeros(self.n_params, np.float32)
        L.call("crh_guide_get_params", self.handle, L.ptr(t), len(t))
        return t

    def eval(self, pos, nrm, wo):
        """Returns (raw outputs (n, 48), mixture (n, 40))."""
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        nrm = np.ascontiguousarray(nrm, np.float32).reshape(-1, 3)
        wo = np.ascontiguousarray(wo, np.float32).reshape(-1, 3)
        n = len(pos)
        raw = np.zeros((n, L.GUIDE_OUTPUTS), np.float32)
        mix = np.zeros((n, 40), np.float32)
        L.call("crh_guide_eval", self.handle, L.ptr(pos), L.ptr(nrm), L.ptr(wo), n, L.ptr(raw),
               L.ptr(mix))
        return raw, mix

    @staticmethod
    def pack_samples(pos, nrm, wo, wi, weight) -> np.ndarray:
        n = len(pos)
        s = np.zeros((n, 16), np.float32)
        s[:, 0:3], s[:, 4:7], s[:, 8:11], s[:, 12:15] = pos, nrm, wo, wi
        s[:, 15] = weight
        return s

    def train(self, samples: np.ndarray, lr: float = 1e-2, apply_adam: bool = True):
        """One step; returns (loss, gradient or None)."""
        s = np.ascontiguousarray(samples, np.float32).reshape(-1, 16)
        loss = np.zeros(1, np.float32)
        grad = None if apply_adam else np.zeros(self.n_params, np.float32)
        L.call("crh_guide_train", self.handle, L.ptr(s), len(s), lr, 1 if apply_adam else 0,
               L.ptr(loss), None if grad is None else L.ptr(grad))
        return float(loss[0]), grad

    def info(self) -> dict:
        """Adam steps taken, samples of the last trained pass, last loss, width."""
        i = L.crh_guide_info()
        L.call("crh_guide_get_info", self.handle, C.byref(i))
        return {"steps": i.steps, "last_samples": i.last_samples, "last_loss": i.last_loss,
                "width": i.width}

    def losses(self, n: int | None = None) -> np.ndarray:
        """Losses of the last n Adam steps (at most 1024), oldest first."""
        steps = self.info()["steps"]
        n = min(steps, 1024) if n is None else n
        out = np.zeros(n, np.float32)
        L.call("crh_guide_get_losses", self.handle, L.ptr(out), n)
        return out

    def __del__(self):
        if getattr(self, "handle", None) and L._lib is not None:
            L._lib.crh_guide_destroy(self.handle)
            self.handle = None


class GuidedIntegrator:
    """Path tracer with one-sample MIS between BSDF sampling (fraction alpha)
    and the guide; optionally trains the guide online on every pass's
    vertices (Adam, KL divergence).  It runs on the fused megakernel (the
    guide evaluated inside the path kernel) or the wavefront, as the
    context's tuning field guided_fused says; both give the same film for
    the same guide."""

    LAMBERTIAN, PRINCIPLED = 1, 2

    def __init__(self, configs, guide: Guide, alpha: float = 0.5, lr: float = 1e-2,
                 train_batch: int = 4096, train_steps: int = 4, max_depth: int = 0,
                 materials: int = 0, target: int = 0):
        """max_depth: guide only vertices at depth < max_depth (0 = all);
        materials: mask of guided BSDFs (LAMBERTIAN | PRINCIPLED, 0 = both);
        target: 0 learns BSDF x cos x incident radiance, 1 incident radiance."""
        self.configs = configs
        self.guide = guide
        self.alpha, self.lr = alpha, lr
        self.train_batch, self.train_steps = train_batch, train_steps
        self.max_depth, self.materials, self.target = max_depth, materials, target
        self.last_stats: dict = {}

    def _cfg(self, train: bool):
        g = L.crh_guided_cfg()
        g.alpha = self.alpha
        g.train = 1 if train else 0
        g.train_batch = self.train_batch
        g.train_steps = self.train_steps
        g.lr = self.lr
        g.max_depth = self.max_depth
        g.materials = self.materials
        g.target = self.target
        return g

    def dispatch(self, scene, spp: int, first_sample: int = 0, train: bool = False,
                 readback: bool = True) -> dict:
        cam = scene.getCamera()
        sensor = cam.getSensor()
        g = self._cfg(train)
        st = L.crh_stats()
        cfg = self.configs.to_c()
        L.call("crh_render_guided", scene.handle, C.byref(cam.data), C.byref(cfg),
               self.guide.handle, C.byref(g), first_sample, spp, sensor.handle, C.byref(st))
        if readback:
            sensor.readbackDeviceData()
        self.last_stats = {"kernel_ms": st.kernel_ms, "main_kernel_ms": st.main_kernel_ms,
                           "total_ms": st.total_ms, "rays": st.rays,
                           "train_samples": st.node_visits}
        return self.last_stats


    def dispatch_views(self, scene, cameras, spp: int, first_sample: int = 0,
                       train: bool = False, readback: bool = True) -> dict:
        """Guided rendering of several camera views of the scene (each with
        its sensor), all of them in each launch, with this guide."""
        sensors = [c.getSensor() for c in cameras]
        if not cameras or any(s is None for s in sensors):
            raise ValueError("every camera needs a sensor")
        cams = (L.crh_camera * len(cameras))(*[c.data for c in cameras])
        films = (C.c_void_p * len(cameras))(*[s.handle.value for s in sensors])
        g = self._cfg(train)
        st = L.crh_stats()
        cfg = self.configs.to_c()
        L.call("crh_render_guided_views", scene.handle, cams, films, len(cameras), C.byref(cfg),
               self.guide.handle, C.byref(g), first_sample, spp, C.byref(st))
        for s in sensors:
            s.setSPP(spp)
            if readback:
                s.readbackDeviceData()
        self.last_stats = {"kernel_ms": st.kernel_ms, "total_ms": st.total_ms, "rays": st.rays,
                           "train_samples": st.node_visits}
        return self.last_stats


def flatten_params(P) -> np.ndarray:
    """oracle.guiding parameter list -> flat crh layout (W row-major, b)."""
    return np.concatenate([np.concatenate([W.reshape(-1), b]) for W, b in P]).astype(np.float32)


def unflatten_params(theta: np.ndarray, width: int = 64):
    shapes = [(64, width), (width, width), (width, width), (width, width), (width, 48)]
    P, off = [], 0
    for k, nn in shapes:
        W = theta[off:off + k * nn].reshape(k, nn).astype(np.float64)
        off += k * nn
        b = theta[off:off + nn].astype(np.float64)
        off += nn
        P.append([W, b])
    return P
