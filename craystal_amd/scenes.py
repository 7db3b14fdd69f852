"""Seeded procedural scenes and ray batches for the benchmark workloads.

Every generator is deterministic: scene layout draws come from SplitMix64
seeded 0xC0FFEE, ray batches from SplitMix64 seeded 0x5EED, both evaluated
as a counter-based stream (value i = mix(seed + (i+1) * golden gamma)).
The same arrays feed the GPU path and the CPU oracle.
"""
from __future__ import annotations

import math

import numpy as np

from .core import (Camera, MaterialData, MaterialFlags, MaterialType, MeshData, SceneData,
                   SphereData)

SCENE_SEED = 0xC0FFEE
RAY_SEED = 0x5EED
_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M = 0xFFFFFFFFFFFFFFFF


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n SplitMix64 outputs of stream `seed`, starting at counter `offset`."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Doubles in [0, 1) from the top 53 bits."""
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


class Stream:
    """Sequential scalar draws from one SplitMix64 stream."""

    def __init__(self, seed: int):
        self.seed, self.i = seed, 0

    def uniform(self, lo=0.0, hi=1.0, n=None):
        k = 1 if n is None else n
        u = uniform01(self.seed, k, self.i)
        self.i += k
        v = lo + (hi - lo) * u
        return float(v[0]) if n is None else v


# ---------------------------------------------------------------------------
# Primitive meshes
# ---------------------------------------------------------------------------
def quad(corner, e1, e2, material, normal=None) -> MeshData:
    """Quad corner, corner+e1, corner+e1+e2, corner+e2 as two triangles with
    counter-clockwise winding around cross(e1, e2)."""
    c = np.asarray(corner, np.float64)
    a, b = np.asarray(e1, np.float64), np.asarray(e2, np.float64)
    P = np.array([c, c + a, c + a + b, c + b], np.float32)
    n = np.cross(a, b) if normal is None else np.asarray(normal, np.float64)
    n = (n / np.linalg.norm(n)).astype(np.float32)
    N = np.tile(n, (4, 1))
    I = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    return MeshData(P, N, I, material, uv)


def box_mesh(center, size, angle_y, material) -> MeshData:
    """Axis-aligned box of `size` rotated about +y by angle_y (radians),
    outward normals, 12 triangles."""
    sx, sy, sz = (np.asarray(size, np.float64) / 2.0)
    ca, sa = math.cos(angle_y), math.sin(angle_y)
    R = np.array([[ca, 0, sa], [0, 1, 0], [-sa, 0, ca]])
    faces = [  # (normal, u-axis, v-axis) in local space, u x v = normal
        ((1, 0, 0), (0, 0, -1), (0, 1, 0)),
        ((-1, 0, 0), (0, 0, 1), (0, 1, 0)),
        ((0, 1, 0), (1, 0, 0), (0, 0, -1)),
        ((0, -1, 0), (1, 0, 0), (0, 0, 1)),
        ((0, 0, 1), (1, 0, 0), (0, 1, 0)),
        ((0, 0, -1), (-1, 0, 0), (0, 1, 0)),
    ]
    half = np.array([sx, sy, sz])
    P, N, I = [], [], []
    for k, (n, u, v) in enumerate(faces):
        n, u, v = map(np.asarray, (n, u, v))
        cen = n * half
        hu = np.abs(u) @ half
        hv = np.abs(v) @ half
        corners = [cen - u * hu - v * hv, cen + u * hu - v * hv, cen + u * hu + v * hv,
                   cen - u * hu + v * hv]
        for q in corners:
            P.append(R @ q + np.asarray(center))
            N.append(R @ n)
        b = 4 * k
        I += [b, b + 1, b + 2, b, b + 2, b + 3]
    return MeshData(np.array(P, np.float32), np.array(N, np.float32), np.array(I, np.uint32),
                    material)


_ICO_CACHE: dict[int, tuple[np.ndarray, np.ndarray]] = {}


def unit_icosphere(subdiv: int):
    """Unit icosphere: 10*4^s + 2 vertices, 20*4^s triangles (float64)."""
    if subdiv in _ICO_CACHE:
        return _ICO_CACHE[subdiv]
    t = (1.0 + math.sqrt(5.0)) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    V = [np.array(v, np.float64) / np.linalg.norm(v) for v in V]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache: dict[tuple[int, int], int] = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = V[a] + V[b]
                V.append(m / np.linalg.norm(m))
                cache[key] = len(V) - 1
            return cache[key]

        nf = []
        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = nf
    out = (np.array(V, np.float64), np.array(F, np.uint32).reshape(-1))
    _ICO_CACHE[subdiv] = out
    return out


def icosphere(center, radius, subdiv, material) -> MeshData:
    V, F = unit_icosphere(subdiv)
    P = (V * radius + np.asarray(center, np.float64)).astype(np.float32)
    return MeshData(P, V.astype(np.float32), F, material)


def _octave_params(rng: "Stream", octaves: int, extent: float, amplitude: float):
    """Per octave (frequency, amplitude, two phases, direction angle): the
    frequency doubles and the amplitude halves from octave to octave."""
    out = []
    for o in range(octaves):
        f = (2.0 ** o) * 2.0 * math.pi / extent * rng.uniform(0.8, 1.6)
        a = amplitude / (2.0 ** o)
        ph1, ph2, ang = rng.uniform(0, 6.283), rng.uniform(0, 6.283), rng.uniform(0, 3.1416)
        out.append((f, a, ph1, ph2, ang))
    return out


def _height(params, X, Z):
    """Height and its x / z derivatives of the octave sum at (X, Z) (float64)."""
    Y = np.zeros_like(X)
    dX, dZ = np.zeros_like(X), np.zeros_like(X)
    for f, a, ph1, ph2, ang in params:
        u = math.cos(ang) * X + math.sin(ang) * Z
        v = -math.sin(ang) * X + math.cos(ang) * Z
        su, cu = np.sin(f * u + ph1), np.cos(f * u + ph1)
        sv, cv = np.sin(f * v + ph2), np.cos(f * v + ph2)
        Y += a * su * cv
        du = a * f * cu * cv
        dv = -a * f * su * sv
        dX += du * math.cos(ang) - dv * math.sin(ang)
        dZ += du * math.sin(ang) + dv * math.cos(ang)
    return Y, dX, dZ


def heightfield(cells: int, extent: float, amplitude: float, seed: int, material,
                octaves: int = 4, params=None) -> MeshData:
    """A (cells x cells)-quad height field over [-extent/2, extent/2]^2 in xz
    (2 cells^2 triangles): a sum of seeded sinusoid octaves (fBm-style:
    frequency x2, amplitude /2 per octave), vertex normals from the analytic
    derivatives.  params: octave parameters from _octave_params (drawn from
    `seed` when None)."""
    if params is None:
        params = _octave_params(Stream(seed), octaves, extent, amplitude)
    n = cells + 1
    t = np.linspace(-extent / 2, extent / 2, n)
    X, Z = np.meshgrid(t, t, indexing="xy")
    Y, dX, dZ = _height(params, X, Z)
    P = np.stack([X, Y, Z], -1).reshape(-1, 3).astype(np.float32)
    del X, Z, Y
    N = np.stack([-dX, np.ones_like(dX), -dZ], -1).reshape(-1, 3)
    del dX, dZ
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    N = N.astype(np.float32)
    i, j = np.meshgrid(np.arange(cells, dtype=np.uint32), np.arange(cells, dtype=np.uint32),
                       indexing="xy")
    a = (j * np.uint32(n) + i).reshape(-1)
    b, c, d = a + 1, a + np.uint32(n + 1), a + np.uint32(n)
    # counter-clockwise seen from +y: (a, d, c), (a, c, b)
    I = np.stack([a, d, c, a, c, b], -1).reshape(-1)
    return MeshData(P, N, I, material)


def principled(base, roughness, metallic) -> MaterialData:
    """A Principled material expressed in the reference's (kd, ks, shininess)
    parameters: roughness = sqrt(2/(n+1)) and metallic = avg(ks)
    (Material/BSDF.cu:21-24, Material/PrincipledBRDF.cu:51-55)."""
    n = 2.0 / (roughness * roughness) - 1.0
    m = MaterialData(diffuseRefl=tuple(base), specularRefl=(metallic,) * 3, shininess=n,
                     type=MaterialType.Principled)
    return m.finalize()


def diffuse(rgb) -> MaterialData:
    return MaterialData(diffuseRefl=tuple(rgb), type=MaterialType.FullDiffuse).finalize()


def phong(kd, ks, n) -> MaterialData:
    return MaterialData(diffuseRefl=tuple(kd), specularRefl=tuple(ks), shininess=n,
                        type=MaterialType.Phong).finalize()


def emitter(radiance) -> MaterialData:
    return MaterialData(emission=tuple(radiance), type=MaterialType.FullDiffuse).finalize()


def _emissive_prims(data: SceneData) -> list[int]:
    """Emissive triangle list (the importer's rule, Utils/Importer.cpp:185-197):
    every triangle whose mesh material is emissive, in primitive-ID order."""
    out, base = [], 0
    for m in data.meshes:
        nt = len(m.index) // 3
        if data.materials[m.materialID].isEmissive():
            out += list(range(base, base + nt))
        base += nt
    return out


# ---------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------
def cornell_box(seed: int = SCENE_SEED, light_radiance=(17.0, 12.0, 4.0)):
    """Cornell-style box: 5 wall quads (10 tris), two boxes rotated about y by
    U[0, 45 deg] (24 tris), a 0.5 x 0.5 m ceiling light (2 tris).  Returns
    (SceneData, camera kwargs)."""
    rng = Stream(seed)
    white, red, green = diffuse((0.73, 0.73, 0.73)), diffuse((0.63, 0.065, 0.05)), \
        diffuse((0.14, 0.45, 0.091))
    light = emitter(light_radiance)
    mats = [white, red, green, light]
    W, R, G, Lm = 0, 1, 2, 3
    meshes = [
        quad((-1, 0, 1), (2, 0, 0), (0, 0, -2), W),      # floor, +y
        quad((-1, 2, -1), (2, 0, 0), (0, 0, 2), W),      # ceiling, -y
        quad((-1, 0, -1), (2, 0, 0), (0, 2, 0), W),      # back wall, +z
        quad((-1, 0, 1), (0, 0, -2), (0, 2, 0), R),      # left wall, +x
        quad((1, 0, -1), (0, 0, 2), (0, 2, 0), G),       # right wall, -x
    ]
    a1 = math.radians(rng.uniform(0.0, 45.0))
    a2 = math.radians(rng.uniform(0.0, 45.0))
    meshes.append(box_mesh((0.35, 0.3, 0.3), (0.6, 0.6, 0.6), a1, W))
    meshes.append(box_mesh((-0.35, 0.6, -0.35), (0.6, 1.2, 0.6), a2, W))
    meshes.append(quad((-0.25, 1.98, -0.25), (0.5, 0, 0), (0, 0, 0.5), Lm))  # faces -y
    data = SceneData(meshes=meshes, materials=mats)
    data.emissiveIndex = _emissive_prims(data)
    cam = dict(pos=(0.0, 1.0, 3.9), target=(0.0, 1.0, 0.0), up=(0.0, 1.0, 0.0), fov=40.0,
               size=(256, 256), spp=16, max_depth=8, rr_start_depth=3)
    return data, cam


def textured_sphere_box(sphere_materials: bool = False, seed: int = SCENE_SEED):
    """The Cornell-style box with three spheres and two textures: an 8 x 8
    RGB checker on the back wall (Lambertian) and a 16 x 4 RGBA gradient on
    the floor (Phong).  sphere_materials=False keeps the reference's rule (a
    sphere takes the material of the triangle with its index: spheres 0
    and 1 the textured floor's, sphere 2 the ceiling's); True gives them a
    textured Principled, the red wall's and the floor's materials.  Returns
    (SceneData, camera kwargs)."""
    data, cam = cornell_box(seed)
    yy, xx = np.mgrid[0:8, 0:8]
    check = np.where(((xx + yy) % 2 == 0)[..., None], (0.85, 0.8, 0.3), (0.15, 0.3, 0.7))
    u = np.linspace(0.0, 1.0, 16, dtype=np.float32)
    grad = np.zeros((4, 16, 4), np.float32)
    grad[..., 0] = u
    grad[..., 1] = 0.6 - 0.4 * u
    grad[..., 2] = np.linspace(0.2, 0.8, 4, dtype=np.float32)[:, None]
    grad[..., 3] = 1.0
    data.textureImages = [check.astype(np.float32), grad]
    tex_flag = MaterialFlags.hasDiffuseTexture
    m4 = MaterialData(diffuseRefl=(0.5, 0.5, 0.5), diffuseHandle=0,
                      type=MaterialType.FullDiffuse).finalize()
    m5 = MaterialData(diffuseRefl=(0.5, 0.5, 0.5), specularRefl=(0.2, 0.2, 0.2), shininess=30.0,
                      diffuseHandle=1, type=MaterialType.Phong).finalize()
    m6 = MaterialData(diffuseRefl=(0.5, 0.5, 0.5), specularRefl=(0.0, 0.0, 0.0), shininess=40.0,
                      diffuseHandle=0, type=MaterialType.Principled).finalize()
    for m in (m4, m5, m6):
        m.flags |= tex_flag
    data.materials += [m4, m5, m6]
    data.meshes[2].materialID = 4      # back wall
    data.meshes[0].materialID = 5      # floor
    data.spheres = [SphereData((-0.45, 0.25, 0.45), 0.25), SphereData((0.45, 0.85, 0.2), 0.22),
                    SphereData((0.05, 1.45, -0.2), 0.18)]
    if sphere_materials:
        data.sphereMaterials = [6, 1, 5]
    return data, cam


def smooth_spectrum(kind: str, scale: float = 1.0) -> np.ndarray:
    """Smooth 95-bin reflectance curves over 360-830 nm for the spectral
    Cornell box: 'white' (nearly flat), 'red' (a logistic edge at 600 nm),
    'green' (a bump at 530 nm), 'flat'."""
    lam = 360.0 + 5.0 * np.arange(95)
    if kind == "white":
        v = 0.73 + 0.03 * np.sin((lam - 360.0) / 470.0 * np.pi)
    elif kind == "red":
        v = 0.05 + 0.6 / (1.0 + np.exp(-(lam - 600.0) / 15.0))
    elif kind == "green":
        v = 0.06 + 0.45 * np.exp(-0.5 * ((lam - 530.0) / 35.0) ** 2)
    elif kind == "flat":
        v = np.ones_like(lam)
    else:
        raise ValueError(kind)
    return (v * scale).astype(np.float32)


def cornell_box_spectral(seed: int = SCENE_SEED, light_radiance: float = 12.0):
    """The Cornell-style box with spectral materials: walls and boxes with
    smooth reflectance curves (white, red, green), a constant-spectrum
    emitter.  The RGB fields hold each spectrum's mean (used by RGB
    transport only).  Returns (SceneData, camera kwargs)."""
    data, cam = cornell_box(seed, light_radiance=(light_radiance,) * 3)
    kinds = ["white", "red", "green"]
    spectra = np.zeros((len(data.materials), 3, 95), np.float32)
    for i, k in enumerate(kinds):
        spectra[i, 0] = smooth_spectrum(k)
        m = float(spectra[i, 0].mean())
        data.materials[i] = diffuse((m, m, m))
    spectra[3, 2] = smooth_spectrum("flat", light_radiance)
    data.spectra = spectra
    return data, cam


def spheres_scene(n_spheres: int = 48, subdiv: int = 5, seed: int = SCENE_SEED,
                  env=(0.05, 0.06, 0.08), light_radiance=(60.0, 60.0, 60.0)):
    """About 1.0M triangles: 48 icospheres (subdivision 5) with radii U[0.2, 1]
    and non-overlapping centres in a 20 x 20 x 5 m box, a 40 x 40 m ground
    quad, a 2 x 2 m area light and a constant environment.  Principled
    materials: baseColor U, roughness U[0.05, 1], metallic Bernoulli(0.3)."""
    rng = Stream(seed)
    mats = [diffuse((0.5, 0.5, 0.5)), emitter(light_radiance)]
    meshes = [quad((-20, 0, 20), (40, 0, 0), (0, 0, -40), 0),
              quad((-1, 10, -1), (2, 0, 0), (0, 0, 2), 1)]
    placed: list[tuple[np.ndarray, float]] = []
    while len(placed) < n_spheres:
        r = rng.uniform(0.2, 1.0)
        c = np.array([rng.uniform(-10, 10), rng.uniform(0, 5), rng.uniform(-10, 10)])
        if any(np.linalg.norm(c - c2) < r + r2 for c2, r2 in placed):
            continue
        placed.append((c, r))
        base = (rng.uniform(), rng.uniform(), rng.uniform())
        rough = rng.uniform(0.05, 1.0)
        metal = 1.0 if rng.uniform() < 0.3 else 0.0
        mats.append(principled(base, rough, metal))
        meshes.append(icosphere(c, r, subdiv, len(mats) - 1))
    data = SceneData(meshes=meshes, materials=mats, envRadiance=env)
    data.emissiveIndex = _emissive_prims(data)
    cam = dict(pos=(0.0, 7.0, 24.0), target=(0.0, 1.5, 0.0), up=(0.0, 1.0, 0.0), fov=45.0,
               size=(1920, 1080), spp=64, max_depth=16, rr_start_depth=3)
    return data, cam


def occluded_room(n_objects: int = 200, subdiv: int = 3, seed: int = SCENE_SEED,
                  light_radiance: float = 50.0, glossy: bool = True):
    """Closed 10 x 10 x 3 m room lit only through a 0.5 x 0.5 m opening in
    the +x wall by an exterior 4 x 4 m area light (radiance 50), furnished
    with seeded boxes and icospheres (about 150k triangles at the default
    sizes) in diffuse and glossy Principled materials (roughness
    U[0.1, 0.6]).  1024 x 1024, max depth 16.  glossy=False gives every
    Principled object a Lambertian material of the same base colour (same
    geometry and random draws)."""
    rng = Stream(seed)
    mats = [diffuse((0.7, 0.7, 0.7)), emitter((light_radiance,) * 3)]
    W = 0
    meshes = [
        quad((-5, 0, 5), (10, 0, 0), (0, 0, -10), W),      # floor, +y
        quad((-5, 3, -5), (10, 0, 0), (0, 0, 10), W),      # ceiling, -y
        quad((-5, 0, -5), (10, 0, 0), (0, 3, 0), W),       # z = -5, +z
        quad((5, 0, 5), (-10, 0, 0), (0, 3, 0), W),        # z = +5, -z
        quad((-5, 0, 5), (0, 0, -10), (0, 3, 0), W),       # x = -5, +x
    ]
    # x = +5 wall (normal -x) with a 0.5 x 0.5 opening at y in [1.25, 1.75],
    # z in [-0.25, 0.25]: four quads around the hole.
    meshes += [
        quad((5, 0, -5), (0, 0, 10), (0, 1.25, 0), W),             # below
        quad((5, 1.75, -5), (0, 0, 10), (0, 1.25, 0), W),          # above
        quad((5, 1.25, -5), (0, 0, 4.75), (0, 0.5, 0), W),         # z < -0.25
        quad((5, 1.25, 0.25), (0, 0, 4.75), (0, 0.5, 0), W),       # z > 0.25
    ]
    # exterior light facing -x, 2 m outside the opening
    meshes.append(quad((7, -0.5, -2), (0, 0, 4), (0, 4, 0), 1))   # faces -x
    for i in range(n_objects):
        x, z = rng.uniform(-4.6, 4.6), rng.uniform(-4.6, 4.6)
        if rng.uniform() < 0.5:
            mats.append(diffuse((rng.uniform(0.3, 0.9), rng.uniform(0.3, 0.9),
                                 rng.uniform(0.3, 0.9))))
        else:
            base = (rng.uniform(), rng.uniform(), rng.uniform())
            rough = rng.uniform(0.1, 0.6)
            mats.append(principled(base, rough, 0.0) if glossy else diffuse(base))
        mid = len(mats) - 1
        if rng.uniform() < 0.5:
            r = rng.uniform(0.15, 0.45)
            sd = subdiv + 1 if rng.uniform() < 0.05 else subdiv
            meshes.append(icosphere((x, r, z), r, sd, mid))
        else:
            sx, sy, sz = rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8)
            meshes.append(box_mesh((x, sy / 2, z), (sx, sy, sz),
                                   math.radians(rng.uniform(0, 90)), mid))
    data = SceneData(meshes=meshes, materials=mats)
    data.emissiveIndex = _emissive_prims(data)
    cam = dict(pos=(-4.2, 1.7, -4.2), target=(2.0, 0.7, 2.0), up=(0.0, 1.0, 0.0), fov=60.0,
               size=(1024, 1024), spp=None, max_depth=16, rr_start_depth=3)
    return data, cam


def batch_scene(index: int, seed: int = SCENE_SEED, tri_range=(1_000, 500_000),
                resolutions=(128, 256, 512, 1024), spps=(4, 16, 64), max_depth: int = 12):
    """One scene of the many-scene batch: a triangle budget drawn
    log-uniformly from tri_range, spent on a height-field ground (about 40%),
    icospheres and boxes; each object gets a Phong (exponent U[1, 1000]) or a
    Principled (roughness U[0.05, 1], metallic Bernoulli(0.3)) material; a
    1.5 x 1.5 m area light and a constant environment.  Square resolution
    and spp drawn from the given sets.  Returns (SceneData, camera kwargs)."""
    rng = Stream(seed * 7919 + 1000 + index)
    lo, hi = math.log(tri_range[0]), math.log(tri_range[1])
    budget = int(math.exp(rng.uniform(lo, hi)))
    mats = [diffuse((0.6, 0.6, 0.6)), emitter((30.0, 28.0, 24.0))]

    def material():
        if rng.uniform() < 0.5:
            kd = (rng.uniform(0.1, 0.8), rng.uniform(0.1, 0.8), rng.uniform(0.1, 0.8))
            ks = rng.uniform(0.05, 0.4)
            mats.append(phong(kd, (ks, ks, ks), rng.uniform(1.0, 1000.0)))
        else:
            base = (rng.uniform(), rng.uniform(), rng.uniform())
            mats.append(principled(base, rng.uniform(0.05, 1.0),
                                   1.0 if rng.uniform() < 0.3 else 0.0))
        return len(mats) - 1

    cells = max(2, int(math.sqrt(0.4 * budget / 2.0)))
    meshes = [heightfield(cells, 12.0, rng.uniform(0.1, 0.8), seed + index, material())]
    left = budget - 2 * cells * cells
    sub = 1
    while sub < 6 and 20 * 4 ** (sub + 1) * 4 <= left:
        sub += 1
    for _ in range(max(1, min(40, left // (20 * 4 ** sub)))):
        r = rng.uniform(0.2, 0.9)
        c = (rng.uniform(-4.5, 4.5), rng.uniform(0.8, 2.5), rng.uniform(-4.5, 4.5))
        meshes.append(icosphere(c, r, sub, material()))
    for _ in range(6):
        size = (rng.uniform(0.3, 1.2), rng.uniform(0.3, 1.5), rng.uniform(0.3, 1.2))
        c = (rng.uniform(-4.5, 4.5), size[1] / 2 + 0.5, rng.uniform(-4.5, 4.5))
        meshes.append(box_mesh(c, size, rng.uniform(0.0, 1.57), material()))
    meshes.append(quad((-0.75, 6.0, -0.75), (1.5, 0, 0), (0, 0, 1.5), 1))  # faces -y
    env = rng.uniform(0.02, 0.1)
    data = SceneData(meshes=meshes, materials=mats, envRadiance=(env, env * 1.05, env * 1.15))
    data.emissiveIndex = _emissive_prims(data)
    res = resolutions[min(len(resolutions) - 1, int(rng.uniform() * len(resolutions)))]
    spp = spps[min(len(spps) - 1, int(rng.uniform() * len(spps)))]
    ang = rng.uniform(0.0, 6.283)
    cam = dict(pos=(9.0 * math.cos(ang), 5.5, 9.0 * math.sin(ang)), target=(0.0, 0.5, 0.0),
               up=(0.0, 1.0, 0.0), fov=50.0, size=(res, res), spp=spp, max_depth=max_depth,
               rr_start_depth=3)
    return data, cam


def disk_emitter(center, normal, radius: float, segments: int, material) -> MeshData:
    """A flat disk as a triangle fan around its centre, facing `normal`
    (counter-clockwise seen from that side; every vertex normal = normal)."""
    nrm = _normalize(np.asarray(normal, np.float64))
    u = _normalize(np.cross(nrm, (0.0, 1.0, 0.0) if abs(nrm[1]) < 0.9 else (1.0, 0.0, 0.0)))
    v = np.cross(nrm, u)
    ang = np.arange(segments) * (2.0 * math.pi / segments)
    rim = (np.asarray(center, np.float64)
           + radius * (np.cos(ang)[:, None] * u + np.sin(ang)[:, None] * v))
    P = np.vstack([np.asarray(center, np.float64)[None], rim]).astype(np.float32)
    N = np.repeat(nrm[None].astype(np.float32), segments + 1, axis=0)
    k = np.arange(segments, dtype=np.uint32)
    I = np.stack([np.zeros_like(k), 1 + k, 1 + (k + 1) % segments], -1).reshape(-1)
    return MeshData(P, N, I, material)


# BVH parameters (bins, max leaf, max depth) of the terrain's fast build:
# the reference's depth cap of 20 (Core/BVH.cu:150-212) leaves the
# 43.8M-triangle terrain with leaves of up to 599 triangles (101.8 triangle
# tests per ray); without it (64 is the GPU builder's limit, the SAH tree
# stops at 29 levels) leaves keep to the 4-triangle rule.  A different tree
# from the reference's, so its hits are checked against the oracle built
# with the same parameters, not against the reference-parameter tree.
TERRAIN_FAST_BVH = (32, 4, 64)


def terrain_scene(cells: int = 4095, extent: float = 1000.0, amplitude: float = 50.0,
                  octaves: int = 8, n_spheres: int = 2000, subdiv: int = 4,
                  size=(3840, 2160), seed: int = SCENE_SEED):
    """The large-scale workload: an fBm height-field terrain ((cells + 1)^2
    vertices, 2 cells^2 triangles; 4095 cells = 33,538,050 triangles, 8
    octaves, amplitude 50 m over 1 km) with n_spheres icospheres (subdivision
    4: 5,120 triangles each) resting on it, glossy Principled materials
    (roughness U[0.02, 0.5]), a small sun-like disk emitter and a constant
    sky.  Four camera views around the centre, 3840 x 2160, 256 spp, max
    depth 24.  Returns (SceneData, [camera kwargs] x 4).  Smaller `cells`,
    `n_spheres`, `subdiv` and `size` give the same layout at test sizes."""
    rng = Stream(seed * 131 + 77)

    def glossy(base, metal=0.0):
        mats.append(principled(base, rng.uniform(0.02, 0.5), metal))
        return len(mats) - 1

    mats = [emitter((2400.0, 2250.0, 2000.0))]
    params = _octave_params(rng, octaves, extent, amplitude)
    meshes = [heightfield(cells, extent, amplitude, 0, glossy((0.32, 0.28, 0.2)), params=params)]
    sun_dir = _normalize(np.array([0.45, 0.62, -0.64]))
    meshes.append(disk_emitter(sun_dir * 0.7 * extent, -sun_dir, 0.02 * extent, 16, 0))
    for _ in range(n_spheres):
        r = rng.uniform(0.002, 0.008) * extent
        x, z = rng.uniform(-0.45, 0.45) * extent, rng.uniform(-0.45, 0.45) * extent
        y = float(_height(params, np.array([x]), np.array([z]))[0][0]) + 0.7 * r
        base = (rng.uniform(), rng.uniform(), rng.uniform())
        metal = 1.0 if rng.uniform() < 0.3 else 0.0
        meshes.append(icosphere((x, y, z), r, subdiv, glossy(base, metal)))
    data = SceneData(meshes=meshes, materials=mats, envRadiance=(0.35, 0.45, 0.65))
    data.emissiveIndex = _emissive_prims(data)
    cams = []
    for k in range(4):
        a = 0.6 + k * math.pi / 2
        px, pz = 0.3 * extent * math.cos(a), 0.3 * extent * math.sin(a)
        tx, tz = 0.05 * extent * math.cos(a + 2.0), 0.05 * extent * math.sin(a + 2.0)
        py = float(_height(params, np.array([px]), np.array([pz]))[0][0]) + 0.06 * extent
        ty = float(_height(params, np.array([tx]), np.array([tz]))[0][0])
        cams.append(dict(pos=(px, py, pz), target=(tx, ty, tz), up=(0.0, 1.0, 0.0), fov=50.0,
                         size=tuple(size), spp=256, max_depth=24, rr_start_depth=3))
    return data, cams


def lampshade_room(light_radiance: float = 40.0):
    """A 4 x 4 x 2.5 m diffuse room lit only indirectly: a 0.4 x 0.4 m
    emitter 0.25 m below the ceiling faces up, a slightly larger diffuse
    shade under it hides its back, so every surface the camera sees is lit
    by the bright ceiling patch around the lamp.  All materials Lambertian.
    Returns (SceneData, camera kwargs)."""
    white, shade = diffuse((0.75, 0.75, 0.75)), diffuse((0.6, 0.6, 0.6))
    mats = [white, shade, emitter((light_radiance,) * 3), diffuse((0.7, 0.3, 0.2))]
    meshes = [
        quad((-2, 0, 2), (4, 0, 0), (0, 0, -4), 0),        # floor, +y
        quad((-2, 2.5, -2), (4, 0, 0), (0, 0, 4), 0),      # ceiling, -y
        quad((-2, 0, -2), (4, 0, 0), (0, 2.5, 0), 0),      # z = -2, +z
        quad((2, 0, 2), (-4, 0, 0), (0, 2.5, 0), 0),       # z = +2, -z
        quad((-2, 0, 2), (0, 0, -4), (0, 2.5, 0), 3),      # x = -2, +x
        quad((2, 0, -2), (0, 0, 4), (0, 2.5, 0), 0),       # x = +2, -x
        quad((-0.2, 2.25, 0.2), (0.4, 0, 0), (0, 0, -0.4), 2),   # lamp, faces +y
        quad((-0.3, 2.24, -0.3), (0.6, 0, 0), (0, 0, 0.6), 1),   # shade, faces -y
        box_mesh((0.6, 0.4, -0.5), (0.8, 0.8, 0.8), 0.5, 0),
    ]
    data = SceneData(meshes=meshes, materials=mats)
    data.emissiveIndex = _emissive_prims(data)
    cam = dict(pos=(-1.6, 1.2, 1.7), target=(0.5, 0.8, -1.0), up=(0.0, 1.0, 0.0), fov=70.0,
               size=(64, 64), spp=16, max_depth=8, rr_start_depth=3)
    return data, cam


def make_camera(cam: dict, sensor=None) -> Camera:
    c = Camera()
    c.setWorldPosition(cam["pos"])
    c.setTarget(cam["target"])
    c.setUp(cam["up"])
    c.setFovY(cam["fov"])
    c.calculateCameraData()
    if sensor is not None:
        c.setSensor(sensor)
    return c


# ---------------------------------------------------------------------------
# Ray batches
# ---------------------------------------------------------------------------
def _normalize(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def offset_ray_origin(p: np.ndarray, n: np.ndarray) -> np.ndarray:
    """offsetRayOrigin (Math/Ray.cu:11-22) on float32 arrays."""
    p = p.astype(np.float32)
    n = n.astype(np.float32)
    of = (np.float32(256.0) * n).astype(np.int32)
    pi = p.view(np.int32) + np.where(p < 0, -of, of)
    pI = pi.astype(np.int32).view(np.float32)
    near = np.abs(p) < np.float32(1.0 / 32.0)
    return np.where(near, p + np.float32(1.0 / 65535.0) * n, pI).astype(np.float32)


def primary_rays(cam: dict, n: int, seed: int = RAY_SEED, offset: int = 0) -> np.ndarray:
    """n camera rays through uniformly drawn pixels + jitter."""
    from . import _lib as L
    W, H = cam["size"]
    u = uniform01(seed, 2 * n, offset).reshape(n, 2)
    px = (u[:, 0] * W).astype(np.float32)
    py = (u[:, 1] * H).astype(np.float32)
    c = make_camera(cam).getData()
    U, V, Wv = (np.array(c.u, np.float32), np.array(c.v, np.float32),
                np.array(c.w, np.float32))
    hh = np.float32(math.tan(math.radians(cam["fov"] * 0.5)))
    hw = np.float32(W) / np.float32(H) * hh
    su = np.float32(2.0) * (px / np.float32(W)) - np.float32(1.0)
    sv = np.float32(1.0) - np.float32(2.0) * (py / np.float32(H))
    d = U[None] * (su * hw)[:, None] + V[None] * (sv * hh)[:, None] + Wv[None]
    d = _normalize(d.astype(np.float64)).astype(np.float32)
    rays = np.zeros(n, L.RAY_DTYPE)
    rays["origin"] = np.array(c.pos, np.float32)
    rays["dir"] = d
    rays["tmin"] = 0.0
    rays["tmax"] = np.inf
    return rays


def secondary_rays(data: SceneData, n: int, seed: int = RAY_SEED, offset: int = 0):
    """n rays leaving surface points: area-weighted triangle pick, uniform
    barycentric point (sampleBarycentric fold), origin offset along the face
    normal, cosine-distributed direction about it."""
    from . import _lib as L
    P, I = data.concatenated()
    tri = I.reshape(-1, 3)
    v0, v1, v2 = (P[tri[:, k]].astype(np.float64) for k in range(3))
    fn = np.cross(v1 - v0, v2 - v0)
    area = 0.5 * np.linalg.norm(fn, axis=1)
    cdf = np.cumsum(area)
    cdf /= cdf[-1]
    u = uniform01(seed, 5 * n, offset).reshape(n, 5)
    t = np.minimum(np.searchsorted(cdf, u[:, 0], side="right"), len(cdf) - 1)
    b1, b2 = u[:, 1], u[:, 2]
    flip = b1 + b2 > 1.0
    b1 = np.where(flip, 1.0 - b1, b1)
    b2 = np.where(flip, 1.0 - b2, b2)
    p = v0[t] * (1 - b1 - b2)[:, None] + v1[t] * b1[:, None] + v2[t] * b2[:, None]
    nrm = _normalize(fn[t])
    r = np.sqrt(u[:, 3])
    phi = 2 * np.pi * u[:, 4]
    lx, ly = r * np.cos(phi), r * np.sin(phi)
    lz = np.sqrt(np.maximum(0.0, 1 - lx * lx - ly * ly))
    a = np.where(np.abs(nrm[:, 1:2]) < 0.9, [[0, 1, 0]], [[1, 0, 0]])
    tx = _normalize(np.cross(a, nrm))
    ty = np.cross(nrm, tx)
    d = _normalize(tx * lx[:, None] + ty * ly[:, None] + nrm * lz[:, None])
    rays = np.zeros(n, L.RAY_DTYPE)
    rays["origin"] = offset_ray_origin(p.astype(np.float32), nrm.astype(np.float32))
    rays["dir"] = d.astype(np.float32)
    rays["tmin"] = 0.0
    rays["tmax"] = np.inf
    return rays


def mixed_batch(data: SceneData, cam: dict, n: int, seed: int = RAY_SEED) -> np.ndarray:
    """n/2 primary + n/2 secondary rays (the closest-hit benchmark batch)."""
    h = n // 2
    return np.concatenate([primary_rays(cam, h, seed, 0),
                           secondary_rays(data, n - h, seed, 2 * h)])
