"""Pins of the oracle's spheres (SphereSOA::intersect, Scene/Sphere.cu:6-49;
SceneView::intersect, Scene/Scene.cu:5-29) and textures
(TextureView::sample, Core/Texture.cu:6-35) against closed forms and brute
force.  All CPU."""
import ctypes as C
import math

import numpy as np
import pytest

import craystal_amd as cr
from craystal_amd import _lib as L
from craystal_amd import scenes


def _ray(o, d, tmin=0.0, tmax=math.inf):
    r = L.crh_ray()
    r.origin[:] = o
    r.dir[:] = d
    r.tmin = tmin
    r.tmax = tmax
    return r


def _sphere_hit(orc, c, rad, o, d, tmin=0.0, tmax=math.inf):
    out = np.zeros(3, np.float32)
    cc = np.asarray(c, np.float32)
    r = _ray(o, d, tmin, tmax)
    ok = orc.lib().orc_sphere_hit(cc.ctypes.data, C.c_float(rad), C.addressof(r), out.ctypes.data)
    return bool(ok), out


def _roots64(c, rad, o, d):
    oc = np.subtract(o, c, dtype=np.float64)
    d = np.asarray(d, np.float64)
    a, b, cc = d @ d, 2 * oc @ d, oc @ oc - rad * rad
    disc = b * b - 4 * a * cc
    if disc < 0:
        return None
    s = math.sqrt(disc)
    return sorted(((-b - s) / (2 * a), (-b + s) / (2 * a)))


def test_sphere_roots_against_float64(orc):
    rng = np.random.default_rng(5)
    n_hit = 0
    for _ in range(3000):
        c = rng.uniform(-2, 2, 3)
        rad = rng.uniform(0.1, 1.5)
        o = rng.uniform(-4, 4, 3)
        d = rng.normal(size=3)
        if rng.uniform() < 0.5:
            d = (c - o) + rng.normal(scale=0.5, size=3)   # aim near the sphere
        d /= np.linalg.norm(d)
        ok, out = _sphere_hit(orc, c, rad, o, d)
        roots = _roots64(c, rad, o, d)
        want = None if roots is None else next((t for t in roots if t >= 0.0), None)
        if want is None or (roots is not None and abs(roots[0] - roots[1]) < 1e-4):
            continue      # misses and grazing tangents: float32 may decide either way
        assert ok, (c, rad, o, d)
        assert abs(out[0] - want) <= 1e-4 * max(1.0, abs(want)), (out[0], want)
        n_hit += 1
    assert n_hit > 800


def test_sphere_inside_range_and_miss_cases(orc):
    # from the centre: the near root is behind the origin, the far one is taken
    ok, out = _sphere_hit(orc, (0, 0, 0), 2.0, (0, 0, 0), (1, 0, 0))
    assert ok and out[0] == 2.0
    # both roots in range: the nearer
    ok, out = _sphere_hit(orc, (5, 0, 0), 1.0, (0, 0, 0), (1, 0, 0))
    assert ok and out[0] == 4.0
    # tMax before the sphere: miss; tMin past the near root: the far root
    assert not _sphere_hit(orc, (5, 0, 0), 1.0, (0, 0, 0), (1, 0, 0), tmax=3.9)[0]
    ok, out = _sphere_hit(orc, (5, 0, 0), 1.0, (0, 0, 0), (1, 0, 0), tmin=4.5)
    assert ok and out[0] == 6.0
    # clear miss, and a sphere behind the ray
    assert not _sphere_hit(orc, (0, 5, 0), 1.0, (0, 0, 0), (1, 0, 0))[0]
    assert not _sphere_hit(orc, (-5, 0, 0), 1.0, (0, 0, 0), (1, 0, 0))[0]


def test_sphere_uv_parameterisation(orc):
    """u = (atan2(-n.z, n.x) + pi) / 2pi, v = acos(-n.y) / pi at the unit
    normal n of the hit point (Scene/Sphere.cu:30-37)."""
    cases = [((0, 5, 0), (0, -1, 0), (None, 1.0)),        # top: n = +y
             ((0, -5, 0), (0, 1, 0), (None, 0.0)),        # bottom: n = -y
             ((5, 0, 0), (-1, 0, 0), (0.5, 0.5)),         # n = +x
             ((-5, 0, 0), (1, 0, 0), (None, 0.5)),        # n = -x: on the seam
             ((0, 0, 5), (0, 0, -1), (0.25, 0.5)),        # n = +z
             ((0, 0, -5), (0, 0, 1), (0.75, 0.5))]        # n = -z
    for o, d, (u, v) in cases:
        ok, out = _sphere_hit(orc, (0, 0, 0), 1.0, o, d)
        assert ok and abs(out[0] - 4.0) < 1e-6
        if u is not None:
            n = -np.asarray(d, np.float64)
            want_u = (math.atan2(-n[2], n[0]) + math.pi) / (2 * math.pi)
            assert abs(out[1] - want_u) < 1e-6, (o, out, want_u)
        assert abs(out[2] - v) < 1e-6, (o, out)


def _sphere_scene(with_mats):
    data, camd = scenes.cornell_box()
    data.spheres = [cr.SphereData((-0.4, 0.3, 0.2), 0.3), cr.SphereData((0.5, 0.8, -0.3), 0.25),
                    cr.SphereData((0.0, 1.2, 0.4), 0.2)]
    if with_mats:
        data.sphereMaterials = [1, 2, 0]
    return data, camd


@pytest.mark.parametrize("with_mats", [False, True])
def test_scene_intersect_spheres_then_bvh_matches_brute_force(orc, with_mats):
    data, camd = _sphere_scene(with_mats)
    osc = orc.Scene(data)
    rays = scenes.mixed_batch(data, camd, 20000)
    hits = osc.trace_closest(rays)
    brute = osc.trace_brute(rays)
    assert np.array_equal(hits["t"], brute["t"])
    same = hits["prim"] == brute["prim"]
    assert same.mean() > 0.999
    sph = (hits["prim"] != L.CRH_INVALID_PRIM) & ((hits["prim"] & L.CRH_HIT_SPHERE) != 0)
    assert sph.sum() > 100
    assert np.all(hits["u"][sph] == 0) and np.all(hits["v"][sph] == 0)
    occ = osc.trace_any(rays)
    assert np.array_equal(occ.astype(bool), hits["prim"] != L.CRH_INVALID_PRIM)


def _tex(w, h, ch, data):
    t = L.crh_texture()
    t.width, t.height, t.channels = w, h, ch
    arr = np.ascontiguousarray(data, np.float32)
    t.data = arr.ctypes.data_as(C.POINTER(C.c_float))
    return t, arr


def _sample(orc, t, u, v):
    out = np.zeros(4, np.float32)
    orc.lib().orc_texture_sample(C.addressof(t), C.c_float(u), C.c_float(v), out.ctypes.data)
    return out


def _bilinear64(img, u, v):
    h, w, c = img.shape
    px, py = u * (w - 1), v * (h - 1)
    x0 = 0 if not px > 0 else int(math.floor(px))
    y0 = 0 if not py > 0 else int(math.floor(py))
    x1, y1 = min(x0 + 1, w - 1), min(y0 + 1, h - 1)
    fx, fy = px - x0, py - y0
    cl = lambda x, y: img[min(y, h - 1), min(x, w - 1)].astype(np.float64)
    c0 = cl(x0, y0) * (1 - fx) + cl(x1, y0) * fx
    c1 = cl(x0, y1) * (1 - fx) + cl(x1, y1) * fx
    return c0 * (1 - fy) + c1 * fy


def test_texture_bilinear_against_closed_form(orc):
    rng = np.random.default_rng(2)
    img = rng.random((5, 7, 4)).astype(np.float32)
    t, keep = _tex(7, 5, 4, img)
    # texel centres are exact
    for y in range(5):
        for x in range(7):
            assert np.array_equal(_sample(orc, t, x / 6, y / 4), img[y, x]), (x, y)
    for u, v in rng.uniform(-0.3, 1.3, (500, 2)):
        want = _bilinear64(img, u, v)
        assert np.allclose(_sample(orc, t, u, v), want, rtol=1e-5, atol=1e-6), (u, v)
    # negative coordinates extrapolate from texel 0 (the cast saturates to 0)
    got = _sample(orc, t, -0.5 / 6, 0.0)
    assert np.allclose(got, img[0, 0] * 1.5 - img[0, 1] * 0.5, rtol=1e-5, atol=1e-6)


def test_texture_channels_fill_with_zero(orc):
    g = np.array([[0.2, 0.4], [0.6, 0.8]], np.float32)
    t, keep = _tex(2, 2, 1, g)
    assert np.array_equal(_sample(orc, t, 1.0, 0.0), np.float32([0.4, 0, 0, 0]))
    rgb = np.arange(12, dtype=np.float32).reshape(2, 2, 3)
    t3, keep3 = _tex(2, 2, 3, rgb)
    assert np.array_equal(_sample(orc, t3, 0.0, 1.0), np.float32([6, 7, 8, 0]))
    one = np.float32([[[0.3, 0.5, 0.7, 1.0]]])
    t1, keep1 = _tex(1, 1, 4, one)
    for u, v in ((0, 0), (0.5, 0.5), (3, -2)):
        assert np.array_equal(_sample(orc, t1, u, v), one[0, 0])
