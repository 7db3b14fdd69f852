"""CPU oracle - TEST INFRASTRUCTURE ONLY.

ctypes wrapper of oracle/_build/liboracle.so, the C++ restatement of the
reference's hot path (oracle/oracle.cpp).  Only tests/, the smoke check and
bench.py's CPU-baseline leg may use it, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        P = C.c_void_p
        u32, f32, u64 = C.c_uint32, C.c_float, C.c_uint64
        sig = {
            "orc_sampler_seq": (None, [u32, u32, u32, u32, P, P]),
            "orc_camera_update": (None, [P]),
            "orc_camera_ray": (None, [P, u32, u32, f32, f32, P]),
            "orc_offset_ray_origin": (None, [P, P, P]),
            "orc_aabb_intersect": (C.c_int, [P, P, P, P]),
            "orc_bvh_build": (P, [P, u32, P, u32, u32, u32, u32]),
            "orc_bvh_stats": (None, [P, P]),
            "orc_bvh_export": (None, [P, P, P]),
            "orc_bvh_free": (None, [P]),
            "orc_scene_create": (P, [P]),
            "orc_cie_cmf": (None, [P]),
            "orc_spectrum_to_xyz": (None, [P, P]),
            "orc_sphere_hit": (C.c_int, [P, f32, P, P]),
            "orc_texture_sample": (None, [P, f32, f32, P]),
            "orc_scene_free": (None, [P]),
            "orc_scene_bvh_stats": (None, [P, P]),
            "orc_scene_bvh_export": (None, [P, P, P]),
            "orc_trace_closest": (None, [P, P, u64, P, P, P, C.c_int]),
            "orc_trace_any": (None, [P, P, u64, P, C.c_int]),
            "orc_trace_brute": (None, [P, P, u64, P]),
            "orc_bsdf_eval_local": (None, [P, P, P, P, P]),
            "orc_bsdf_sample_local": (None, [P, P, P, P, P, P]),
            "orc_film_splat": (None, [P, P, u32, u32, P, f32, f32]),
            "orc_render_samples": (u64, [P, P, u32, u32, P, u32, P, P, C.c_int]),
            "orc_render_samples_counted": (u64, [P, P, u32, u32, P, u32, P, P, C.c_int, P]),
            "orc_render": (u64, [P, P, u32, u32, P, u32, u32, P, P, C.c_int]),
            "orc_max_threads": (C.c_int, []),
            "orc_render_guided": (u64, [P, P, u32, u32, P, P, f32, u32, u32, C.c_int, u32, u32,
                                        P, P, C.c_int]),
            "orc_render_samples_guided": (u64, [P, P, u32, u32, P, P, f32, u32, u32, C.c_int,
                                                u32, P, P, C.c_int]),
            "orc_mixture": (None, [P, P]),
            "orc_mixture_pdf": (f32, [P, P]),
            "orc_mixture_sample": (None, [P, P, P]),
            "orc_actual_pdf_local": (f32, [P, P, P]),
            "orc_guided_sample_local": (None, [P, P, f32, P, P, P, P, P]),
            "orc_bsdf_sample_dirs": (None, [P, P, P, u64, P]),
            "orc_actual_pdf_batch": (None, [P, P, P, u64, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
    return _lib


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _ref(s):
    return C.cast(C.byref(s), C.c_void_p)


class Scene:
    """Oracle scene built from a craystal_amd SceneData."""

    def __init__(self, data):
        desc, keep = data.to_c()
        self._keep = keep
        self.h = lib().orc_scene_create(_ref(desc))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_scene_free(self.h)
            self.h = None

    def bvh_stats(self):
        from craystal_amd import _lib as L
        st = L.crh_bvh_stats()
        lib().orc_scene_bvh_stats(self.h, _ref(st))
        return {"node_count": st.node_count, "max_depth": st.max_depth,
                "max_leaf": st.max_leaf, "root": st.root, "prim_count": st.prim_count,
                "internal_count": st.internal_count, "build_ms": st.build_ms}

    def export_bvh(self):
        from craystal_amd import _lib as L
        st = self.bvh_stats()
        nodes = np.zeros(st["node_count"], L.NODE_DTYPE)
        prims = np.zeros(st["prim_count"], np.uint32)
        lib().orc_scene_bvh_export(self.h, _p(nodes), _p(prims))
        return nodes, prims

    def trace_closest(self, rays, counts=False, threads=0):
        from craystal_amd import _lib as L
        rays = np.ascontiguousarray(rays, dtype=L.RAY_DTYPE)
        hits = np.zeros(len(rays), L.HIT_DTYPE)
        nv = np.zeros(len(rays), np.uint32) if counts else None
        nt = np.zeros(len(rays), np.uint32) if counts else None
        lib().orc_trace_closest(self.h, _p(rays), len(rays), _p(hits), _p(nv), _p(nt), threads)
        return (hits, nv, nt) if counts else hits

    def trace_any(self, rays, threads=0):
        from craystal_amd import _lib as L
        rays = np.ascontiguousarray(rays, dtype=L.RAY_DTYPE)
        occ = np.zeros(len(rays), np.uint8)
        lib().orc_trace_any(self.h, _p(rays), len(rays), _p(occ), threads)
        return occ.astype(bool)

    def trace_brute(self, rays):
        from craystal_amd import _lib as L
        rays = np.ascontiguousarray(rays, dtype=L.RAY_DTYPE)
        hits = np.zeros(len(rays), L.HIT_DTYPE)
        lib().orc_trace_brute(self.h, _p(rays), len(rays), _p(hits))
        return hits

    def render_samples(self, cam, size, cfg, sample_index, threads=0, counts=False):
        """One sample index of every pixel: (radiance, film positions, rays
        traced); counts=True adds the MIS BSDF rays among them that cannot hit
        their sampled emitter triangle (a renderer may skip those)."""
        W, H = size
        Lr = np.zeros((H, W, 3), np.float32)
        P = np.zeros((H, W, 2), np.float32)
        c = cfg.to_c()
        if counts:
            un = C.c_uint64(0)
            rays = lib().orc_render_samples_counted(self.h, _ref(cam.getData()), W, H, _ref(c),
                                                    sample_index, _p(Lr), _p(P), threads,
                                                    C.byref(un))
            return Lr, P, rays, un.value
        rays = lib().orc_render_samples(self.h, _ref(cam.getData()), W, H, _ref(c),
                                        sample_index, _p(Lr), _p(P), threads)
        return Lr, P, rays

    def render_guided(self, cam, size, cfg, out48, first, spp, alpha=0.5, mask=3,
                      max_depth=0, bsdf_as_guide=False, threads=0):
        """Guided film: out48 = the guide's 48 outputs used at every vertex
        (None with bsdf_as_guide: the proposal is the BSDF's own sampler)."""
        W, H = size
        s = np.zeros((H, W, 3), np.float32)
        w = np.zeros((H, W), np.float32)
        c = cfg.to_c()
        o = None if out48 is None else np.ascontiguousarray(out48, np.float32)
        rays = lib().orc_render_guided(self.h, _ref(cam.getData()), W, H, _ref(c), _p(o),
                                       alpha, mask, max_depth, 1 if bsdf_as_guide else 0,
                                       first, spp, _p(s), _p(w), threads)
        return s, w, rays

    def render_samples_guided(self, cam, size, cfg, out48, sample_index, alpha=0.5, mask=3,
                              max_depth=0, bsdf_as_guide=False, threads=0):
        W, H = size
        Lr = np.zeros((H, W, 3), np.float32)
        P = np.zeros((H, W, 2), np.float32)
        c = cfg.to_c()
        o = None if out48 is None else np.ascontiguousarray(out48, np.float32)
        rays = lib().orc_render_samples_guided(self.h, _ref(cam.getData()), W, H, _ref(c), _p(o),
                                               alpha, mask, max_depth,
                                               1 if bsdf_as_guide else 0, sample_index, _p(Lr),
                                               _p(P), threads)
        return Lr, P, rays

    def render(self, cam, size, cfg, first, spp, threads=0):
        W, H = size
        s = np.zeros((H, W, 3), np.float32)
        w = np.zeros((H, W), np.float32)
        c = cfg.to_c()
        rays = lib().orc_render(self.h, _ref(cam.getData()), W, H, _ref(c), first, spp,
                                _p(s), _p(w), threads)
        return s, w, rays


def build_bvh(pos, idx, bins=32, max_leaf=4, max_depth=20):
    """Reference BVH build on a triangle soup -> (nodes, prim_order, stats)."""
    from craystal_amd import _lib as L
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    idx = np.ascontiguousarray(idx, np.uint32).reshape(-1)
    h = lib().orc_bvh_build(_p(pos), len(pos), _p(idx), len(idx) // 3, bins, max_leaf, max_depth)
    st = L.crh_bvh_stats()
    lib().orc_bvh_stats(h, _ref(st))
    nodes = np.zeros(st.node_count, L.NODE_DTYPE)
    prims = np.zeros(st.prim_count, np.uint32)
    lib().orc_bvh_export(h, _p(nodes), _p(prims))
    lib().orc_bvh_free(h)
    stats = {"node_count": st.node_count, "max_depth": st.max_depth, "max_leaf": st.max_leaf,
             "root": st.root, "root_min": list(st.root_min), "root_max": list(st.root_max),
             "build_ms": st.build_ms}
    return nodes, prims, stats


def sampler_seq(x, y, idx, n):
    out = np.zeros(n, np.float32)
    raw = np.zeros(n, np.uint32)
    lib().orc_sampler_seq(x, y, idx, n, _p(out), _p(raw))
    return out, raw


def camera_ray(cam_c, W, H, px, py):
    from craystal_amd import _lib as L
    r = L.crh_ray()
    lib().orc_camera_ray(_ref(cam_c), W, H, px, py, _ref(r))
    return np.array(r.origin), np.array(r.dir), r.tmin, r.tmax


def camera_update(cam_c):
    lib().orc_camera_update(_ref(cam_c))


def offset_ray_origin(p, n):
    p = np.ascontiguousarray(p, np.float32)
    n = np.ascontiguousarray(n, np.float32)
    out = np.zeros(3, np.float32)
    lib().orc_offset_ray_origin(_p(p), _p(n), _p(out))
    return out


def aabb_intersect(pmin, pmax, ray):
    from craystal_amd import _lib as L
    pmin = np.ascontiguousarray(pmin, np.float32)
    pmax = np.ascontiguousarray(pmax, np.float32)
    r = np.ascontiguousarray(ray, dtype=L.RAY_DTYPE).reshape(1)
    t = np.zeros(1, np.float32)
    h = lib().orc_aabb_intersect(_p(pmin), _p(pmax), _p(r), _p(t))
    return bool(h), float(t[0])


def bsdf_eval_local(mat, wo, wi):
    m = mat.to_c()
    wo = np.ascontiguousarray(wo, np.float32)
    wi = np.ascontiguousarray(wi, np.float32)
    f = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    lib().orc_bsdf_eval_local(_ref(m), _p(wo), _p(wi), _p(f), _p(pdf))
    return f, float(pdf[0])


def bsdf_sample_local(mat, wo, u):
    m = mat.to_c()
    wo = np.ascontiguousarray(wo, np.float32)
    u = np.ascontiguousarray(u, np.float32)
    wi = np.zeros(3, np.float32)
    f = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    lib().orc_bsdf_sample_local(_ref(m), _p(wo), _p(u), _p(wi), _p(f), _p(pdf))
    return wi, f, float(pdf[0])


def film_splat(sum_, wsum, L3, px, py):
    H, W = wsum.shape
    L3 = np.ascontiguousarray(L3, np.float32)
    lib().orc_film_splat(_p(sum_), _p(wsum), W, H, _p(L3), px, py)


def mixture(out48):
    """(weights (8), means (8, 3), kappas (8)) of the guide head."""
    o = np.ascontiguousarray(out48, np.float32)
    m = np.zeros(40, np.float32)
    lib().orc_mixture(_p(o), _p(m))
    return m[:8], m[8:32].reshape(8, 3), m[32:40]


def mixture_pdf(out48, d):
    o = np.ascontiguousarray(out48, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    return float(lib().orc_mixture_pdf(_p(o), _p(d)))


def mixture_sample(out48, u3):
    o = np.ascontiguousarray(out48, np.float32)
    u = np.ascontiguousarray(u3, np.float32)
    d = np.zeros(3, np.float32)
    lib().orc_mixture_sample(_p(o), _p(u), _p(d))
    return d


def actual_pdf_local(mat, wo, wi):
    m = mat.to_c()
    wo = np.ascontiguousarray(wo, np.float32)
    wi = np.ascontiguousarray(wi, np.float32)
    return float(lib().orc_actual_pdf_local(_ref(m), _p(wo), _p(wi)))


def guided_sample_local(mat, out48, alpha, wo, u4):
    """One guided continuation (identity frame) from the uniforms u4; out48
    None = the BSDF as the proposal.  Returns (wi, f * |cos|, pdf)."""
    m = mat.to_c()
    o = None if out48 is None else np.ascontiguousarray(out48, np.float32)
    wo = np.ascontiguousarray(wo, np.float32)
    u = np.ascontiguousarray(u4, np.float32)
    wi = np.zeros(3, np.float32)
    f = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    lib().orc_guided_sample_local(_ref(m), _p(o), alpha, _p(wo), _p(u), _p(wi), _p(f), _p(pdf))
    return wi, f, float(pdf[0])


def bsdf_sample_dirs(mat, wo, u):
    """Directions of BSDF::sampleLocal for the uniforms u (n, 2)."""
    m = mat.to_c()
    wo = np.ascontiguousarray(wo, np.float32)
    u = np.ascontiguousarray(u, np.float32).reshape(-1, 2)
    wi = np.zeros((len(u), 3), np.float32)
    lib().orc_bsdf_sample_dirs(_ref(m), _p(wo), _p(u), len(u), _p(wi))
    return wi


def actual_pdf_batch(mat, wo, wi):
    m = mat.to_c()
    wo = np.ascontiguousarray(wo, np.float32)
    wi = np.ascontiguousarray(wi, np.float32).reshape(-1, 3)
    out = np.zeros(len(wi), np.float32)
    lib().orc_actual_pdf_batch(_ref(m), _p(wo), _p(wi), len(wi), _p(out))
    return out
