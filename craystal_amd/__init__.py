"""craystal_amd - MI355X-native rendering hot path of the CRaystal path tracer.

gfx950 HIP kernels behind a C-ABI (include/crayhip.h, libcrayhip.so) and a
host-side mirror of the reference's Scene / Camera / Sensor /
PathTraceIntegrator API.  Build the library with
``python -m craystal_amd.build``.
"""
from ._lib import RuntimeError, load  # noqa: A004
from .core import (BVH, SPECTRUM_BINS, SPECTRUM_LAMBDA, Camera, Configs, DeviceBuffer,
                   MaterialData, MaterialFlags, MaterialType, MeshData, PathTraceIntegrator,
                   Scene, SceneData, Sensor, SphereData, cie_cmf, device_count, get_context,
                   rgb_to_spectrum, spectrum_to_xyz, trace_any_device, trace_closest_device)

__all__ = ["RuntimeError", "load", "BVH", "Camera", "Configs", "DeviceBuffer", "MaterialData",
           "MaterialFlags", "MaterialType", "MeshData", "PathTraceIntegrator", "Scene",
           "SceneData", "SphereData", "Sensor", "device_count", "get_context", "trace_any_device",
           "trace_closest_device", "SPECTRUM_BINS", "SPECTRUM_LAMBDA", "cie_cmf",
           "spectrum_to_xyz", "rgb_to_spectrum"]
