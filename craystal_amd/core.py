"""Host-side mirror of the reference's rendering API over the C-ABI.

Same names and argument meaning as the reference's C++ classes:
``MaterialData`` / ``MeshData`` / ``SceneData`` (Core/Material.h:34-66,
Scene/Triangle.h:36-42, Scene/Scene.h:40-47), ``Scene`` (Scene/Scene.h:49-77),
``Camera`` (Core/Camera.h:12-50), ``Sensor`` (Core/Sensor.h:31-70),
``PathTraceIntegrator`` (Core/Integrator.h:10-38) and ``BVH``
(Core/BVH.h:48-61).  Device memory is owned by libcrayhip.so; errors raise
:class:`craystal_amd.RuntimeError`.  There is no CPU fallback: if the library
is missing, constructing any device object raises ImportError.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

SPECTRUM_BINS = 95          # 5 nm bins from 360 nm to 830 nm
SPECTRUM_LAMBDA = 360.0 + 5.0 * np.arange(SPECTRUM_BINS)


def cie_cmf() -> np.ndarray:
    """(95, 3) x, y, z-bar used by spectral transport (crh_cie_cmf)."""
    out = np.zeros((SPECTRUM_BINS, 3), np.float32)
    L.call("crh_cie_cmf", L.ptr(out))
    return out


def spectrum_to_xyz(spectrum) -> np.ndarray:
    """XYZ of a 95-bin spectrum; a constant 1 gives Y = 1."""
    s = np.ascontiguousarray(spectrum, np.float32).reshape(SPECTRUM_BINS)
    out = np.zeros(3, np.float32)
    L.call("crh_spectrum_to_xyz", L.ptr(s), L.ptr(out))
    return out


def rgb_to_spectrum(rgb) -> np.ndarray:
    """95-bin spectrum of an RGB value (crh_rgb_to_spectrum: Smits's
    decomposition; the replacement of BroadSpectrum::fromRGB,
    Core/Spectrum.cpp:49).  The white-balanced linear sRGB of the result
    equals rgb; gray (k, k, k) gives the flat spectrum k."""
    c = np.ascontiguousarray(rgb, np.float32).reshape(3)
    out = np.zeros(SPECTRUM_BINS, np.float32)
    L.call("crh_rgb_to_spectrum", L.ptr(c), L.ptr(out))
    return out


# sRGB -> XYZ (Core/SpectrumTabData.h:63-65), used for the XYZ film output.
SRGB_TO_XYZ = np.array([[0.412453, 0.357580, 0.180423],
                        [0.212671, 0.715160, 0.072169],
                        [0.019334, 0.119193, 0.950227]], dtype=np.float32)

_contexts: dict[int, "Context"] = {}


class Context:
    """One device context (stream + events) per device index."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        L.call("crh_context_create", device, C.byref(h))
        self.handle = h
        self.device = device

    TUNING_FIELDS = tuple(k for k, _ in L.crh_tuning._fields_)

    def set_tuning(self, **kw):
        """Launch choices that never change results (see crh_tuning)."""
        t = L.crh_tuning()
        for k, v in kw.items():
            if k not in self.TUNING_FIELDS:
                raise KeyError(k)
            setattr(t, k, int(v))
        L.call("crh_context_set_tuning", self.handle, C.byref(t))

    def get_tuning(self) -> dict:
        t = L.crh_tuning()
        L.call("crh_context_get_tuning", self.handle, C.byref(t))
        return {k: getattr(t, k) for k in self.TUNING_FIELDS}

    def __del__(self):
        if getattr(self, "handle", None) and L._lib is not None:
            L._lib.crh_context_destroy(self.handle)
            self.handle = None


def get_context(device: int = 0) -> Context:
    if device not in _contexts:
        _contexts[device] = Context(device)
    return _contexts[device]


def device_count() -> int:
    n = C.c_int(0)
    L.call("crh_device_count", C.byref(n))
    return n.value


class MaterialType:
    FullDiffuse = L.MAT_DIFFUSE
    Principled = L.MAT_PRINCIPLED
    Phong = L.MAT_PHONG


class MaterialFlags:
    IsEmissive = L.MATF_EMISSIVE
    IsDiffuse = L.MATF_DIFFUSE
    hasDiffuseTexture = L.MATF_DIFFUSE_TEXTURE


@dataclass
class MaterialData:
    diffuseRefl: tuple = (0.0, 0.0, 0.0)
    diffuseHandle: int = 0
    specularRefl: tuple = (0.0, 0.0, 0.0)
    transmittance: tuple = (0.0, 0.0, 0.0)
    emission: tuple = (0.0, 0.0, 0.0)
    shininess: float = 0.0
    IoR: float = 1.0
    type: int = MaterialType.FullDiffuse
    flags: int = 0

    def finalize(self) -> "MaterialData":
        """MaterialData::finalize (Core/Material.h:57-65)."""
        m = self.to_c()
        L.call("crh_material_finalize", C.byref(m))
        self.flags = m.flags
        return self

    def isEmissive(self) -> bool:
        return bool(self.flags & MaterialFlags.IsEmissive)

    def to_c(self) -> L.crh_material:
        m = L.crh_material()
        m.diffuse[:] = [float(x) for x in self.diffuseRefl]
        m.diffuse_tex = self.diffuseHandle
        m.specular[:] = [float(x) for x in self.specularRefl]
        m.transmittance[:] = [float(x) for x in self.transmittance]
        m.emission[:] = [float(x) for x in self.emission]
        m.shininess = self.shininess
        m.ior = self.IoR
        m.type = self.type
        m.flags = self.flags
        return m


@dataclass
class MeshData:
    position: np.ndarray
    normal: np.ndarray
    index: np.ndarray
    materialID: int = 0
    texCrd: np.ndarray | None = None

    def __post_init__(self):
        self.position = np.ascontiguousarray(self.position, dtype=np.float32).reshape(-1, 3)
        self.normal = np.ascontiguousarray(self.normal, dtype=np.float32).reshape(-1, 3)
        self.index = np.ascontiguousarray(self.index, dtype=np.uint32).reshape(-1)
        if self.texCrd is not None:
            self.texCrd = np.ascontiguousarray(self.texCrd, dtype=np.float32).reshape(-1, 2)


@dataclass
class SphereData:
    """SphereData (Scene/Sphere.h)."""
    center: tuple = (0.0, 0.0, 0.0)
    radius: float = 1.0


@dataclass
class SceneData:
    """SceneData (Scene/Scene.h:40-47): meshes, spheres, materials, texture
    images ((H, W, C) float arrays, C = 1..4) and the emissive triangle list,
    plus this library's environment, BVH parameters and spectra."""
    meshes: list = field(default_factory=list)
    materials: list = field(default_factory=list)
    emissiveIndex: list = field(default_factory=list)
    spheres: list = field(default_factory=list)
    textureImages: list = field(default_factory=list)
    # Per-sphere material ids; None = the reference's rule (a sphere takes the
    # material of the triangle with its index).
    sphereMaterials: object = None
    # Constant environment radiance on a miss (zero = the reference).
    envRadiance: tuple = (0.0, 0.0, 0.0)
    # BVH parameters (bins, max leaf, max depth); zeros take the reference's.
    bvhParams: tuple = (32, 4, 20)
    bvhDevice: bool = False  # build the BVH on the GPU (bit-identical to the host build)
    # Spectral transport: (nmaterial, 3, 95) diffuse, specular and emission
    # spectra, 5 nm bins from 360 nm; the environment's 95 bins.  None: the
    # RGB values are uplifted (rgb_to_spectrum) when the scene is created.
    spectra: object = None
    envSpectrum: object = None

    @property
    def triangle_count(self) -> int:
        return sum(len(m.index) // 3 for m in self.meshes)

    def concatenated(self):
        """Global (positions, indices) in primitive-ID order."""
        pos, idx, off = [], [], 0
        for m in self.meshes:
            pos.append(m.position)
            idx.append(m.index.astype(np.uint64) + off)
            off += len(m.position)
        P = np.concatenate(pos) if pos else np.zeros((0, 3), np.float32)
        I = np.concatenate(idx).astype(np.uint32) if idx else np.zeros(0, np.uint32)
        return np.ascontiguousarray(P), np.ascontiguousarray(I)

    def to_c(self):
        """Build a crh_scene_desc; returns (desc, keepalive)."""
        keep = []
        meshes = (L.crh_mesh * max(1, len(self.meshes)))()
        for i, m in enumerate(self.meshes):
            keep += [m.position, m.normal, m.index]
            meshes[i].position = m.position.ctypes.data_as(C.POINTER(C.c_float))
            meshes[i].normal = m.normal.ctypes.data_as(C.POINTER(C.c_float))
            if m.texCrd is not None:
                keep.append(m.texCrd)
                meshes[i].texcrd = m.texCrd.ctypes.data_as(C.POINTER(C.c_float))
            meshes[i].nvert = len(m.position)
            meshes[i].index = m.index.ctypes.data_as(C.POINTER(C.c_uint32))
            meshes[i].nindex = len(m.index)
            meshes[i].material_id = m.materialID
        mats = (L.crh_material * max(1, len(self.materials)))()
        for i, m in enumerate(self.materials):
            mats[i] = m.to_c()
        em = np.ascontiguousarray(self.emissiveIndex, dtype=np.uint32)
        keep += [meshes, mats, em]
        d = L.crh_scene_desc()
        d.meshes = C.cast(meshes, C.POINTER(L.crh_mesh))
        d.nmesh = len(self.meshes)
        d.materials = C.cast(mats, C.POINTER(L.crh_material))
        d.nmaterial = len(self.materials)
        d.emissive_index = em.ctypes.data_as(C.POINTER(C.c_uint32))
        d.nemissive = len(em)
        d.env_radiance[:] = [float(x) for x in self.envRadiance]
        d.bvh.bins, d.bvh.max_leaf, d.bvh.max_depth = self.bvhParams[:3]
        d.bvh.device = 1 if self.bvhDevice else 0
        if self.spectra is not None:
            sp = np.ascontiguousarray(self.spectra, dtype=np.float32)
            if sp.shape != (len(self.materials), 3, SPECTRUM_BINS):
                raise ValueError(f"spectra must be (nmaterial, 3, {SPECTRUM_BINS})")
            keep.append(sp)
            d.spectra = sp.ctypes.data_as(C.POINTER(C.c_float))
            if self.envSpectrum is not None:
                env = np.ascontiguousarray(self.envSpectrum, dtype=np.float32).reshape(-1)
                if env.shape != (SPECTRUM_BINS,):
                    raise ValueError(f"envSpectrum must have {SPECTRUM_BINS} bins")
                keep.append(env)
                d.env_spectrum = env.ctypes.data_as(C.POINTER(C.c_float))
        if self.spheres:
            cen = np.ascontiguousarray([sp.center for sp in self.spheres], np.float32)
            rad = np.ascontiguousarray([sp.radius for sp in self.spheres], np.float32)
            keep += [cen, rad]
            d.sphere_center = cen.ctypes.data_as(C.POINTER(C.c_float))
            d.sphere_radius = rad.ctypes.data_as(C.POINTER(C.c_float))
            d.nsphere = len(self.spheres)
            if self.sphereMaterials is not None:
                sm = np.ascontiguousarray(self.sphereMaterials, np.uint32).reshape(-1)
                if len(sm) != len(self.spheres):
                    raise ValueError("sphereMaterials needs one id per sphere")
                keep.append(sm)
                d.sphere_material = sm.ctypes.data_as(C.POINTER(C.c_uint32))
        if self.textureImages:
            tex = (L.crh_texture * len(self.textureImages))()
            for i, im in enumerate(self.textureImages):
                a = np.asarray(im, np.float32)
                if a.ndim == 2:
                    a = a[..., None]
                a = np.ascontiguousarray(a)
                keep.append(a)
                tex[i].height, tex[i].width, tex[i].channels = a.shape
                tex[i].data = a.ctypes.data_as(C.POINTER(C.c_float))
            keep.append(tex)
            d.textures = C.cast(tex, C.POINTER(L.crh_texture))
            d.ntexture = len(self.textureImages)
        return d, keep


def _stats_dict(st: L.crh_bvh_stats) -> dict:
    return {"node_count": st.node_count, "max_depth": st.max_depth, "max_leaf": st.max_leaf,
            "root": st.root, "prim_count": st.prim_count,
            "internal_count": st.internal_count, "root_min": list(st.root_min),
            "root_max": list(st.root_max), "build_ms": st.build_ms}


class BVH:
    """BVH::build (Core/BVH.cu:88-338) on a triangle soup: on the host, or
    with device=True on the GPU (crh_bvh_build_device, the same result bit
    for bit)."""

    def __init__(self, positions: np.ndarray, indices: np.ndarray, bins=32, max_leaf=4,
                 max_depth=20, device: bool = False, gpu: int = 0):
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
        p = L.crh_bvh_params(bins, max_leaf, max_depth, 0)
        h = C.c_void_p()
        if device:
            L.call("crh_bvh_build_device", get_context(gpu).handle, L.ptr(pos), len(pos),
                   L.ptr(idx), len(idx) // 3, C.byref(p), C.byref(h))
        else:
            L.call("crh_bvh_build", L.ptr(pos), len(pos), L.ptr(idx), len(idx) // 3,
                   C.byref(p), C.byref(h))
        self.handle = h

    def stats(self) -> dict:
        st = L.crh_bvh_stats()
        L.call("crh_bvh_get_stats", self.handle, C.byref(st))
        return _stats_dict(st)

    def export(self):
        st = self.stats()
        nodes = np.zeros(st["node_count"], L.NODE_DTYPE)
        prims = np.zeros(st["prim_count"], np.uint32)
        L.call("crh_bvh_export", self.handle, L.ptr(nodes), L.ptr(prims))
        return nodes, prims

    def __del__(self):
        if getattr(self, "handle", None) and L._lib is not None:
            L._lib.crh_bvh_destroy(self.handle)
            self.handle = None


class Sensor:
    """Film of tent-filtered spectrum sums and weight sums (Core/Sensor.h:31-70)."""

    def __init__(self, size, spp: int = 1, device: int = 0):
        self.size = (int(size[0]), int(size[1]))
        self.spp = int(spp)
        self.ctx = get_context(device)
        h = C.c_void_p()
        L.call("crh_film_create", self.ctx.handle, self.size[0], self.size[1], C.byref(h))
        self.handle = h
        self.data = np.zeros((self.size[1], self.size[0], 3), np.float32)
        self.weight = np.zeros((self.size[1], self.size[0]), np.float32)

    def setSPP(self, v: int):
        self.spp = int(v)

    def getSPP(self) -> int:
        return self.spp

    def getSize(self):
        return self.size

    def clear(self):
        L.call("crh_film_clear", self.handle)

    def readbackDeviceData(self):
        L.call("crh_film_readback", self.handle, L.ptr(self.data), L.ptr(self.weight))

    def createImage(self) -> np.ndarray:
        """Linear RGB image = sum / weight (Core/Sensor.cu:93-112)."""
        with np.errstate(divide="ignore", invalid="ignore"):
            return self.data / self.weight[..., None]

    def createImageXYZ(self) -> np.ndarray:
        """CIE XYZ fp32 image of the same film."""
        return self.createImage() @ SRGB_TO_XYZ.T

    def __del__(self):
        if getattr(self, "handle", None) and L._lib is not None:
            L._lib.crh_film_destroy(self.handle)
            self.handle = None


class Camera:
    """Look-at pinhole camera (Core/Camera.h:12-50, Core/CameraProxy.h)."""

    def __init__(self):
        self.data = L.crh_camera()
        self.data.pos[:] = [0.0, 0.0, 0.0]
        self.data.fov_y = 60.0
        self.data.target[:] = [0.0, 0.0, -1.0]
        self.data.near_z = 0.01
        self.data.up[:] = [0.0, 1.0, 0.0]
        self.data.far_z = 10000.0
        self.data.u[:] = [0.0, 0.0, 1.0]
        self.data.v[:] = [0.0, 1.0, 0.0]
        self.data.w[:] = [1.0, 0.0, 0.0]
        self.sensor: Sensor | None = None

    def setFovY(self, v):
        self.data.fov_y = float(v)

    def setWorldPosition(self, p):
        self.data.pos[:] = [float(x) for x in p]

    def setTarget(self, p):
        self.data.target[:] = [float(x) for x in p]

    def setUp(self, p):
        self.data.up[:] = [float(x) for x in p]

    def setSensor(self, s: Sensor):
        self.sensor = s

    def getSensor(self) -> Sensor | None:
        return self.sensor

    def calculateCameraData(self):
        L.call("crh_camera_update", C.byref(self.data))

    def getData(self) -> L.crh_camera:
        return self.data


class Scene:
    """Builds the BVH and uploads geometry/materials (Scene/Scene.cu:61-78)."""

    def __init__(self, data: SceneData, device: int = 0):
        self.ctx = get_context(device)
        desc, keep = data.to_c()
        h = C.c_void_p()
        L.call("crh_scene_create", self.ctx.handle, C.byref(desc), C.byref(h))
        self.handle = h
        self.camera: Camera | None = None
        self.triangle_count = data.triangle_count

    def getCamera(self) -> Camera | None:
        return self.camera

    def setCamera(self, cam: Camera):
        self.camera = cam

    def bvh_stats(self) -> dict:
        st = L.crh_bvh_stats()
        L.call("crh_scene_get_bvh_stats", self.handle, C.byref(st))
        return _stats_dict(st)

    def export_bvh(self):
        st = self.bvh_stats()
        nodes = np.zeros(st["node_count"], L.NODE_DTYPE)
        prims = np.zeros(st["prim_count"], np.uint32)
        L.call("crh_scene_export_bvh", self.handle, L.ptr(nodes), L.ptr(prims))
        return nodes, prims

    def trace_closest(self, rays: np.ndarray) -> np.ndarray:
        """SceneView::intersect over a host ray batch (Scene/Scene.cu:5-14)."""
        rays = np.ascontiguousarray(rays, dtype=L.RAY_DTYPE)
        hits = np.zeros(len(rays), L.HIT_DTYPE)
        L.call("crh_trace_closest", self.handle, L.ptr(rays), len(rays), L.ptr(hits))
        return hits

    def trace_any(self, rays: np.ndarray) -> np.ndarray:
        """SceneView::intersectOcclusion (Scene/Scene.cu:16-29)."""
        rays = np.ascontiguousarray(rays, dtype=L.RAY_DTYPE)
        occ = np.zeros(len(rays), np.uint8)
        L.call("crh_trace_any", self.handle, L.ptr(rays), len(rays), L.ptr(occ))
        return occ.astype(bool)

    def __del__(self):
        if getattr(self, "handle", None) and L._lib is not None:
            L._lib.crh_scene_destroy(self.handle)
            self.handle = None


class DeviceBuffer:
    """DeviceBuffer (Core/Buffer.h:13-44) with size_t sizes."""

    def __init__(self, nbytes: int, device: int = 0):
        self.ctx = get_context(device)
        h = C.c_void_p()
        L.call("crh_buffer_create", self.ctx.handle, int(nbytes), C.byref(h))
        self.handle = h
        self.nbytes = int(nbytes)

    def copyFromHost(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        L.call("crh_buffer_upload", self.handle, L.ptr(a), a.nbytes)

    def copyToHost(self, a: np.ndarray):
        L.call("crh_buffer_download", self.handle, L.ptr(a), a.nbytes)

    def memset(self, value: int = 0):
        L.call("crh_buffer_memset", self.handle, value)

    def copyFrom(self, src: "DeviceBuffer", nbytes: int | None = None):
        """Device-to-device copy."""
        L.call("crh_buffer_copy", self.handle, src.handle,
               int(min(self.nbytes, src.nbytes) if nbytes is None else nbytes))

    def size(self) -> int:
        n = C.c_size_t()
        L.call("crh_buffer_size", self.handle, C.byref(n))
        return n.value

    def __del__(self):
        if getattr(self, "handle", None) and L._lib is not None:
            L._lib.crh_buffer_destroy(self.handle)
            self.handle = None


@dataclass
class Configs:
    """PathTraceIntegrator::Configs (Core/Integrator.h:14-18), plumbed to the
    device.  The reference's kernel always ran with maxDepth 10,
    rrThreshold 0.05 and RR when depth > 5: see :meth:`reference`."""
    maxDepth: int = 10
    rrThreshold: float = 0.05
    useMIS: bool = True
    rrStartDepth: int = 6
    mode: int = L.MODE_MEGAKERNEL
    # Spectral transport: three of the 95 wavelength bins per path; the film
    # then holds CIE XYZ (spectra from SceneData.spectra or uplifted RGB).
    spectral: bool = False

    @staticmethod
    def reference(useMIS: bool = True) -> "Configs":
        return Configs(10, 0.05, useMIS, 6)

    def to_c(self) -> L.crh_integrator_cfg:
        c = L.crh_integrator_cfg()
        c.max_depth = self.maxDepth
        c.rr_threshold = self.rrThreshold
        c.rr_start_depth = self.rrStartDepth
        c.use_mis = 1 if self.useMIS else 0
        c.mode = self.mode
        c.spectral = 1 if self.spectral else 0
        return c


class PathTraceIntegrator:
    """PathTraceIntegrator::dispatch (Core/Integrator.cu:233-262)."""

    def __init__(self, configs: Configs | None = None):
        self.configs = configs or Configs()
        self.last_stats: dict = {}

    def dispatch(self, scene: Scene, spp: int, first_sample: int = 0,
                 readback: bool = True) -> dict:
        cam = scene.getCamera()
        if cam is None or cam.getSensor() is None:
            raise ValueError("scene camera and sensor must be set before dispatch")
        sensor = cam.getSensor()
        sensor.setSPP(spp)
        st = L.crh_stats()
        cfg = self.configs.to_c()
        L.call("crh_render", scene.handle, C.byref(cam.data), C.byref(cfg), first_sample, spp,
               sensor.handle, C.byref(st))
        if readback:
            sensor.readbackDeviceData()
        self.last_stats = {"kernel_ms": st.kernel_ms, "main_kernel_ms": st.main_kernel_ms,
                           "total_ms": st.total_ms, "rays": st.rays, "launches": st.launches}
        return self.last_stats


def trace_closest_device(scene: Scene, rays: DeviceBuffer, n: int, hits: DeviceBuffer,
                         count_work: bool = False) -> dict:
    st = L.crh_stats()
    L.call("crh_trace_closest_device", scene.handle, rays.handle, n, hits.handle,
           1 if count_work else 0, C.byref(st))
    return dataclasses.asdict(_Stats(st.kernel_ms, st.total_ms, st.rays, st.node_visits,
                                     st.tri_tests, st.launches))


def trace_any_device(scene: Scene, rays: DeviceBuffer, n: int, occ: DeviceBuffer) -> dict:
    st = L.crh_stats()
    L.call("crh_trace_any_device", scene.handle, rays.handle, n, occ.handle, C.byref(st))
    return {"kernel_ms": st.kernel_ms, "total_ms": st.total_ms, "rays": st.rays}


@dataclass
class _Stats:
    kernel_ms: float
    total_ms: float
    rays: int
    node_visits: int
    tri_tests: int
    launches: int
