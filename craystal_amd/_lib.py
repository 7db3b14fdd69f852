"""ctypes binding of libcrayhip.so (include/crayhip.h).

Struct layouts mirror the header exactly; numpy dtypes are provided for the
array types (rays, hits, BVH nodes) so batches cross the boundary without
copies.  Nonzero status codes raise :class:`RuntimeError` carrying the
library's message (the reference's ``CRay::RuntimeError``, Core/Error.h:59-66).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("CRH_LIBRARY", PKG / "libcrayhip.so"))

u32 = C.c_uint32
f32 = C.c_float

CRH_OK = 0
CRH_ERR_INVALID_ARG = 1
CRH_ERR_HIP = 2
CRH_ERR_OOM = 3
CRH_ERR_UNSUPPORTED = 4
CRH_ERR_INTERNAL = 5
CRH_INVALID_PRIM = 0xFFFFFFFF
CRH_HIT_SPHERE = 0x40000000   # crh_hit.prim of a sphere hit: index | this bit

MAT_DIFFUSE, MAT_PRINCIPLED, MAT_PHONG = 0, 1, 2
MATF_EMISSIVE, MATF_DIFFUSE, MATF_DIFFUSE_TEXTURE = 1, 2, 4
MODE_MEGAKERNEL, MODE_WAVEFRONT = 0, 1


class RuntimeError(Exception):  # noqa: A001 - mirrors CRay::RuntimeError
    """Raised for every nonzero status of the C-ABI."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class crh_ray(C.Structure):
    _fields_ = [("origin", f32 * 3), ("tmin", f32), ("dir", f32 * 3), ("tmax", f32)]


class crh_hit(C.Structure):
    _fields_ = [("prim", u32), ("t", f32), ("u", f32), ("v", f32)]


class crh_bvh_params(C.Structure):
    _fields_ = [("bins", u32), ("max_leaf", u32), ("max_depth", u32), ("device", u32)]


class crh_bvh_node_ref(C.Structure):
    _fields_ = [("tag", u32), ("a", u32), ("b", u32), ("pmin", f32 * 3), ("pmax", f32 * 3)]


class crh_bvh_stats(C.Structure):
    _fields_ = [("node_count", u32), ("max_depth", u32), ("max_leaf", u32), ("root", u32),
                ("prim_count", u32), ("internal_count", u32), ("root_min", f32 * 3),
                ("root_max", f32 * 3), ("build_ms", C.c_double)]


class crh_material(C.Structure):
    _fields_ = [("diffuse", f32 * 3), ("diffuse_tex", u32), ("specular", f32 * 3),
                ("transmittance", f32 * 3), ("emission", f32 * 3), ("shininess", f32),
                ("ior", f32), ("type", u32), ("flags", u32)]


class crh_mesh(C.Structure):
    _fields_ = [("position", C.POINTER(f32)), ("normal", C.POINTER(f32)),
                ("texcrd", C.POINTER(f32)), ("nvert", u32), ("index", C.POINTER(u32)),
                ("nindex", u32), ("material_id", u32)]


class crh_texture(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("channels", u32), ("data", C.POINTER(f32))]


class crh_scene_desc(C.Structure):
    _fields_ = [("meshes", C.POINTER(crh_mesh)), ("nmesh", u32),
                ("materials", C.POINTER(crh_material)), ("nmaterial", u32),
                ("emissive_index", C.POINTER(u32)), ("nemissive", u32),
                ("env_radiance", f32 * 3), ("bvh", crh_bvh_params),
                ("spectra", C.POINTER(f32)), ("env_spectrum", C.POINTER(f32)),
                ("sphere_center", C.POINTER(f32)), ("sphere_radius", C.POINTER(f32)),
                ("nsphere", u32), ("sphere_material", C.POINTER(u32)),
                ("textures", C.POINTER(crh_texture)), ("ntexture", u32)]


class crh_camera(C.Structure):
    _fields_ = [("pos", f32 * 3), ("fov_y", f32), ("target", f32 * 3), ("near_z", f32),
                ("up", f32 * 3), ("far_z", f32), ("u", f32 * 3), ("pad0", u32),
                ("v", f32 * 3), ("pad1", u32), ("w", f32 * 3), ("pad2", u32)]


class crh_integrator_cfg(C.Structure):
    _fields_ = [("max_depth", u32), ("rr_threshold", f32), ("rr_start_depth", u32),
                ("use_mis", u32), ("mode", u32), ("count_work", u32),
                ("spectral", u32), ("reserved", u32)]


class crh_tuning(C.Structure):
    _fields_ = [(k, u32) for k in ("stack", "trace_block", "persistent", "refill", "pt_grid",
                                   "render_stack", "render_pool", "render_waves", "render_regen",
                                   "render_grid", "wf_finish", "wf_stack", "wf_trace_grid", "wf_trace_block",
                                   "wf_shade_grid", "wf_persistent", "wf_refill", "wf_pool",
                                   "batch_stack", "batch_tile_w", "batch_grid", "batch_chunks",
                                   "batch_graph", "batch_waves", "render_band",
                                   "batch_slots", "trace_tris", "guided_fused", "guided_stack",
                                   "guided_waves", "guided_grid", "guided_pool")]


class crh_batch_info(C.Structure):
    _fields_ = [("scenes", u32), ("max_stack", u32), ("pool_bytes", C.c_uint64),
                ("nodes", C.c_uint64), ("triangles", C.c_uint64)]


class crh_guided_cfg(C.Structure):
    _fields_ = [("alpha", f32), ("train", u32), ("train_batch", u32), ("train_steps", u32),
                ("lr", f32), ("max_depth", u32), ("materials", u32), ("target", u32)]


class crh_guide_info(C.Structure):
    _fields_ = [("steps", u32), ("last_samples", u32), ("last_loss", f32), ("width", u32)]


class crh_stats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("main_kernel_ms", C.c_double),
                ("total_ms", C.c_double), ("rays", C.c_uint64),
                ("node_visits", C.c_uint64), ("tri_tests", C.c_uint64),
                ("launches", C.c_uint64)]


RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("tmin", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("prim", "<u4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])
NODE_DTYPE = np.dtype([("tag", "<u4"), ("a", "<u4"), ("b", "<u4"), ("pmin", "<f4", 3),
                       ("pmax", "<f4", 3)])
assert RAY_DTYPE.itemsize == C.sizeof(crh_ray) == 32
assert HIT_DTYPE.itemsize == C.sizeof(crh_hit) == 16
assert NODE_DTYPE.itemsize == C.sizeof(crh_bvh_node_ref) == 36
assert C.sizeof(crh_material) == 68 and C.sizeof(crh_camera) == 96

_P = C.c_void_p
# name -> (argtypes); every function returns int.
SIGNATURES = {
    "crh_device_count": [C.POINTER(C.c_int)],
    "crh_version": [C.POINTER(u32)],
    "crh_context_create": [C.c_int, C.POINTER(_P)],
    "crh_context_destroy": [_P],
    "crh_context_set_tuning": [_P, C.POINTER(crh_tuning)],
    "crh_context_get_tuning": [_P, C.POINTER(crh_tuning)],
    "crh_buffer_create": [_P, C.c_size_t, C.POINTER(_P)],
    "crh_buffer_upload": [_P, _P, C.c_size_t],
    "crh_buffer_download": [_P, _P, C.c_size_t],
    "crh_buffer_memset": [_P, C.c_int],
    "crh_buffer_size": [_P, C.POINTER(C.c_size_t)],
    "crh_buffer_copy": [_P, _P, C.c_size_t],
    "crh_buffer_destroy": [_P],
    "crh_bvh_build": [_P, u32, _P, u32, C.POINTER(crh_bvh_params), C.POINTER(_P)],
    "crh_bvh_get_stats": [_P, C.POINTER(crh_bvh_stats)],
    "crh_bvh_export": [_P, _P, _P],
    "crh_bvh_destroy": [_P],
    "crh_material_finalize": [C.POINTER(crh_material)],
    "crh_scene_create": [_P, C.POINTER(crh_scene_desc), C.POINTER(_P)],
    "crh_scene_get_bvh_stats": [_P, C.POINTER(crh_bvh_stats)],
    "crh_scene_export_bvh": [_P, _P, _P],
    "crh_scene_destroy": [_P],
    "crh_camera_update": [C.POINTER(crh_camera)],
    "crh_trace_closest": [_P, _P, C.c_uint64, _P],
    "crh_trace_any": [_P, _P, C.c_uint64, _P],
    "crh_trace_closest_device": [_P, _P, C.c_uint64, _P, C.c_int, C.POINTER(crh_stats)],
    "crh_trace_any_device": [_P, _P, C.c_uint64, _P, C.POINTER(crh_stats)],
    "crh_bvh_build_device": [_P, _P, u32, _P, u32, C.POINTER(crh_bvh_params), C.POINTER(_P)],
    "crh_trace_closest_device_noalloc": [_P, _P, C.c_uint64, _P, C.c_int,
                                         C.POINTER(crh_stats)],
    "crh_film_create": [_P, u32, u32, C.POINTER(_P)],
    "crh_film_clear": [_P],
    "crh_film_readback": [_P, _P, _P],
    "crh_film_destroy": [_P],
    "crh_render": [_P, C.POINTER(crh_camera), C.POINTER(crh_integrator_cfg), u32, u32, _P,
                   C.POINTER(crh_stats)],
    "crh_guide_create": [_P, _P, _P, u32, C.POINTER(_P)],
    "crh_guide_create_width": [_P, _P, _P, u32, u32, C.POINTER(_P)],
    "crh_guide_set_params": [_P, _P, u32],
    "crh_guide_get_params": [_P, _P, u32],
    "crh_guide_eval": [_P, _P, _P, _P, u32, _P, _P],
    "crh_guide_train": [_P, _P, u32, f32, C.c_int, _P, _P],
    "crh_guide_get_info": [_P, C.POINTER(crh_guide_info)],
    "crh_guide_get_losses": [_P, _P, u32],
    "crh_guide_destroy": [_P],
    "crh_cie_cmf": [_P],
    "crh_spectrum_to_xyz": [_P, _P],
    "crh_rgb_to_spectrum": [_P, _P],
    "crh_batch_create": [_P, _P, u32, _P],
    "crh_batch_get_info": [_P, C.POINTER(crh_batch_info)],
    "crh_batch_render": [_P, _P, _P, _P, C.POINTER(crh_integrator_cfg), u32,
                         C.POINTER(crh_stats)],
    "crh_batch_destroy": [_P],
    "crh_render_guided": [_P, C.POINTER(crh_camera), C.POINTER(crh_integrator_cfg), _P,
                          C.POINTER(crh_guided_cfg), u32, u32, _P, C.POINTER(crh_stats)],
    "crh_render_guided_views": [_P, _P, _P, u32, C.POINTER(crh_integrator_cfg), _P,
                                C.POINTER(crh_guided_cfg), u32, u32, C.POINTER(crh_stats)],
}
GUIDE_PARAMS = 19760
GUIDE_PARAMS_128 = 64048
GUIDE_OUTPUTS = 48

_lib = None


def load() -> C.CDLL:
    """Load libcrayhip.so; raises ImportError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: run `python -m craystal_amd.build`")
    lib = C.CDLL(str(LIB_PATH))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.crh_last_error.argtypes = []
    lib.crh_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != CRH_OK:
        msg = load().crh_last_error().decode(errors="replace")
        raise RuntimeError(status, msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))


def ptr(a: np.ndarray) -> C.c_void_p:
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)
