"""Pins of the CPU oracle (oracle/oracle.cpp) against what fixes its answers
independently of its own code: the reference's own tests, closed forms,
textbook special cases and brute force.  All CPU.
"""
import math

import numpy as np
import pytest

import craystal_amd as cr
from craystal_amd import _lib as L
from craystal_amd import scenes


# ---------------------------------------------------------------------------
# Sampler (Core/Sampler.cu:4-31) - ports of Tests/CoreTest.cpp:130-281
# ---------------------------------------------------------------------------
def test_sampler_reference_properties(orc):
    s, _ = orc.sampler_seq(123, 456, 0, 10000)
    assert (s >= 0).all() and (s < 1).all()                       # :134-142
    h = np.histogram(s, bins=10, range=(0, 1))[0]                 # :172-190
    assert ((h > 800) & (h < 1200)).all()
    assert abs(s.mean() - 0.5) < 0.05                              # :250-260
    assert abs(s.var(ddof=1) - 1 / 12) < 0.01                      # :262-280
    a, _ = orc.sampler_seq(0, 0, 0, 100)
    b, _ = orc.sampler_seq(0, 1, 0, 100)
    assert (a != b).any()                                          # :158-170
    c, _ = orc.sampler_seq(123, 456, 1, 100)
    assert (c != s[:100]).any()                                    # :232-246
    d, _ = orc.sampler_seq(123, 456, 0, 100)
    assert np.array_equal(d, s[:100])                              # :222-230
    xy = s[:2000].reshape(-1, 2)                                   # :192-214
    q = [(xy[:, 0] >= .5) & (xy[:, 1] >= .5), (xy[:, 0] < .5) & (xy[:, 1] >= .5),
         (xy[:, 0] < .5) & (xy[:, 1] < .5), (xy[:, 0] >= .5) & (xy[:, 1] < .5)]
    assert all(x.any() for x in q)


def test_sampler_known_answers(orc):
    """The stream is fully specified integer arithmetic; these constants are
    worked by hand from the recurrence for pixel (0, 0), sample 0:
    state0 = 0, inc = 1; four warm-up steps give state 0x8B68F8D4 ... the
    output of step k is RXS-M-XS of the state *before* the step."""
    M = 0xFFFFFFFF

    def rxsmxs(o):
        o = (((o >> ((o >> 28) + 4)) ^ o) * 277803737) & M
        return (o >> 22) ^ o

    st, inc = 0, 1
    for _ in range(4):
        st = (st * 747796405 + inc) & M
    expect = []
    for _ in range(5):
        expect.append(rxsmxs(st))
        st = (st * 747796405 + inc) & M
    _, raw = orc.sampler_seq(0, 0, 0, 5)
    assert list(raw) == expect
    # stateToFloat maps u32 >= 2^32 - 128 to exactly 1.0f (float rounding)
    assert np.float32(np.float32(2 ** 32 - 100) * np.float32(1.0 / 4294967296.0)) == 1.0


# ---------------------------------------------------------------------------
# Camera (Core/Camera.cu:35-39, Core/CameraProxy.h:25-40) - Tests/CameraTest.cpp
# ---------------------------------------------------------------------------
CAMERA_KATS = [((256, 256), (0.0, 0.0, -1.0), 1e-5),
               ((0, 0), (-0.44721359, 0.44721359, -0.774596691), 1e-4),
               ((511, 511), (0.446162969, -0.446162969, -0.77580744), 1e-4)]


@pytest.mark.parametrize("pos,expect,tol", CAMERA_KATS)
def test_camera_reference_kats(orc, pos, expect, tol):
    cam = cr.Camera()                       # default: fov 60, origin, looking -z
    orc.camera_update(cam.data)
    o, d, tmin, tmax = orc.camera_ray(cam.data, 512, 512, *pos)
    assert np.array_equal(o, [0, 0, 0])
    assert np.allclose(d, expect, atol=tol)
    assert tmin == 0.0 and math.isinf(tmax)


def test_camera_update_product_equals_oracle(orc):
    for p, t in [((0, 1, 3.9), (0, 1, 0)), ((3, 2, -5), (0.3, 0.1, 0.2)), ((0, 7, 24), (0, 1.5, 0))]:
        a = cr.Camera()
        a.setWorldPosition(p)
        a.setTarget(t)
        b = cr.Camera()
        b.setWorldPosition(p)
        b.setTarget(t)
        a.calculateCameraData()          # product (libcrayhip, host code)
        orc.camera_update(b.data)        # oracle
        for f in ("u", "v", "w"):
            assert np.array_equal(np.array(getattr(a.data, f), np.float32),
                                  np.array(getattr(b.data, f), np.float32))


# ---------------------------------------------------------------------------
# offsetRayOrigin (Math/Ray.cu:11-22), Ray Tracing Gems ch. 6
# ---------------------------------------------------------------------------
def test_offset_ray_origin_kats(orc):
    # |p| < 1/32 uses p + n/65535; else integer-ulp nudge of 256*n ulps
    out = orc.offset_ray_origin([1.0, 0.01, -2.0], [0.0, 0.0, 1.0])
    assert out[0] == np.float32(1.0)                              # n.x = 0 -> no nudge
    assert out[1] == np.float32(0.01)                             # near-origin, n.y = 0
    assert out[2] == np.float32(-2.0 + 256 * 2.0 ** -23)          # 256 ulps toward +z
    out = orc.offset_ray_origin([0.5, -0.5, 0.0], [1.0, -1.0, 1.0])
    assert out[0] == np.float32(0.5 + 256 * 2.0 ** -24)
    assert out[1] == np.float32(-0.5 - 256 * 2.0 ** -24)
    assert out[2] == np.float32(np.float32(1.0 / 65535.0))
    # truncation of 256 * n toward zero: n = 0.5/256 gives no int offset
    out = orc.offset_ray_origin([4.0, 4.0, 4.0], [0.5 / 256, -0.9 / 256, 1.9 / 256])
    assert out[0] == np.float32(4.0) and out[1] == np.float32(4.0)
    assert out[2] == np.float32(4.0 + 2.0 ** -21)


def test_offset_ray_origin_numpy_matches(orc):
    rng = np.random.default_rng(1)
    p = rng.normal(size=(200, 3)).astype(np.float32) * np.float32(3)
    p[:20] *= np.float32(0.005)
    n = rng.normal(size=(200, 3))
    n = (n / np.linalg.norm(n, axis=1, keepdims=True)).astype(np.float32)
    vec = scenes.offset_ray_origin(p, n)
    for i in range(200):
        assert np.array_equal(vec[i], orc.offset_ray_origin(p[i], n[i]))


# ---------------------------------------------------------------------------
# AABB slab test (Math/AABB.cu:6-46)
# ---------------------------------------------------------------------------
def _ray(o, d, tmin=0.0, tmax=np.inf):
    r = np.zeros(1, L.RAY_DTYPE)
    r["origin"], r["dir"], r["tmin"], r["tmax"] = o, d, tmin, tmax
    return r


def test_aabb_special_cases(orc):
    lo, hi = [0, 0, 0], [1, 1, 1]
    assert orc.aabb_intersect(lo, hi, _ray([-2, .5, .5], [1, 0, 0])) == (True, 2.0)
    assert orc.aabb_intersect(lo, hi, _ray([.5, .5, .5], [0, 0, 1])) == (True, 0.0)
    assert not orc.aabb_intersect(lo, hi, _ray([-2, .5, .5], [-1, 0, 0]))[0]
    assert not orc.aabb_intersect(lo, hi, _ray([-2, 1.5, .5], [1, 0, 0]))[0]
    assert not orc.aabb_intersect(lo, hi, _ray([-2, .5, .5], [1, 0, 0], 0, 1.5))[0]
    # flat box (z extent 0): plane hit uses the division branch
    flo, fhi = [0, 0, 2], [1, 1, 2]
    hit, t = orc.aabb_intersect(flo, fhi, _ray([.5, .5, 0], [0, 0, 1]))
    assert hit and t == 2.0
    # flat box, ray parallel inside the plane (|o - min| <= eps): accepted
    assert orc.aabb_intersect(flo, fhi, _ray([-1, .5, 2], [1, 0, 0]))[0]
    # flat box, ray parallel off the plane: rejected
    assert not orc.aabb_intersect(flo, fhi, _ray([-1, .5, 2.01], [1, 0, 0]))[0]
    # axis-parallel ray exactly on a slab plane: (min-o)*inf = NaN is ignored
    # by the std::max/std::min selects
    assert orc.aabb_intersect(lo, hi, _ray([0, .5, -1], [0, 0, 1]))[0]


def test_aabb_vs_float64_slab(orc):
    rng = np.random.default_rng(3)
    for _ in range(400):
        a, b = rng.uniform(-2, 2, 3), rng.uniform(-2, 2, 3)
        lo, hi = np.minimum(a, b).astype(np.float32), np.maximum(a, b).astype(np.float32)
        o = rng.uniform(-4, 4, 3).astype(np.float32)
        d = rng.normal(size=3).astype(np.float32)
        t0 = (lo.astype(np.float64) - o) / d
        t1 = (hi.astype(np.float64) - o) / d
        tn = max(0.0, np.minimum(t0, t1).max())
        tf = np.maximum(t0, t1).min()
        if abs(tn - tf) < 1e-4:
            continue  # grazing: float32 rounding decides
        hit, t = orc.aabb_intersect(lo, hi, _ray(o, d))
        assert hit == (tn <= tf)
        if hit:
            assert abs(t - tn) <= 1e-4 * max(1.0, abs(tn))


# ---------------------------------------------------------------------------
# Traversal + Moller-Trumbore vs brute force (Core/BVH.cu:13-84,
# Scene/Triangle.cu:40-75, Scene/Shape.h:17-37)
# ---------------------------------------------------------------------------
def _sphere_scene(params=(32, 4, 20)):
    mat = scenes.diffuse((0.5, 0.5, 0.5))
    return cr.SceneData(meshes=[scenes.icosphere((0, 0, 0), 1.0, 3, 0),
                                scenes.icosphere((1.2, 0.3, -0.4), 0.6, 2, 0),
                                scenes.quad((-3, -1, 3), (6, 0, 0), (0, 0, -6), 0)],
                        materials=[mat], bvhParams=params)


@pytest.mark.parametrize("params", [(32, 4, 20), (4, 1, 30), (32, 8, 3)])
def test_traversal_equals_brute_force(orc, params):
    data = _sphere_scene(params)
    s = orc.Scene(data)
    cam = dict(pos=(0.3, 1.0, 4.0), target=(0, 0, 0), up=(0, 1, 0), fov=60, size=(64, 64))
    rays = scenes.mixed_batch(data, cam, 3000)
    h = s.trace_closest(rays)
    b = s.trace_brute(rays)
    assert np.array_equal(h["t"].view(np.uint32), b["t"].view(np.uint32))
    same = h["prim"] == b["prim"]
    # a differing ID is only allowed on an exact t tie between two triangles
    for i in np.nonzero(~same)[0]:
        assert h["t"][i] == b["t"][i]
    assert same.mean() > 0.99
    assert s.trace_any(rays).tolist() == (h["prim"] != L.CRH_INVALID_PRIM).tolist()


def test_moller_trumbore_vs_float64(orc):
    """Single-triangle scenes: hit/miss and t against a float64 solve."""
    rng = np.random.default_rng(5)
    mat = scenes.diffuse((1, 1, 1))
    for _ in range(60):
        P = rng.uniform(-1, 1, (3, 3)).astype(np.float32)
        data = cr.SceneData(meshes=[cr.MeshData(P, np.tile([0, 0, 1], (3, 1)), [0, 1, 2], 0)],
                            materials=[mat])
        s = orc.Scene(data)
        rays = np.zeros(50, L.RAY_DTYPE)
        o = rng.uniform(-2, 2, (50, 3)).astype(np.float32)
        tgt = (P[0] + rng.uniform(-0.5, 1.5, (50, 1)) * (P[1] - P[0])
               + rng.uniform(-0.5, 1.5, (50, 1)) * (P[2] - P[0]))
        d = (tgt - o).astype(np.float32)
        rays["origin"], rays["dir"], rays["tmax"] = o, d, np.inf
        h = s.trace_closest(rays)
        e1, e2 = (P[1] - P[0]).astype(np.float64), (P[2] - P[0]).astype(np.float64)
        for i in range(50):
            A = np.stack([-d[i].astype(np.float64), e1, e2], 1)
            if abs(np.linalg.det(A)) < 1e-5:
                continue
            t, u, v = np.linalg.solve(A, o[i].astype(np.float64) - P[0])
            margin = min(u, v, 1 - u - v)
            if abs(margin) < 1e-4 or abs(t) < 1e-4:
                continue
            inside = margin > 0 and t > 0
            assert (h["prim"][i] == 0) == inside
            if inside:
                assert abs(h["t"][i] - t) <= 1e-4 * max(1, t)
                assert abs(h["u"][i] - u) < 1e-4 and abs(h["v"][i] - v) < 1e-4


# ---------------------------------------------------------------------------
# BVH build (Core/BVH.cu:88-338): invariants and the SAH selection rule
# ---------------------------------------------------------------------------
def _check_bvh(nodes, prims, P, I, bins=32, max_leaf=4, max_depth=20):
    n = len(nodes)
    tri = I.reshape(-1, 3)
    V = P[tri]                                   # (ntri, 3, 3)
    cen = ((V[:, 0] + V[:, 1]) + V[:, 2]) / np.float32(3.0)
    assert sorted(prims.tolist()) == list(range(len(tri)))
    root = n - 1
    depth = {root: 0}
    seen_leaf = np.zeros(len(tri), bool)
    for i in range(n - 1, -1, -1):
        nd = nodes[i]
        if nd["tag"] == 1:
            for c in (nd["a"], nd["b"]):
                assert c < i                      # post-order: children first
                depth[c] = depth[i] + 1
                assert (nodes[c]["pmin"] >= nd["pmin"]).all()
                assert (nodes[c]["pmax"] <= nd["pmax"]).all()
        else:
            off, cnt = nd["a"], nd["b"]
            seen_leaf[prims[off:off + cnt]] = True
            tb = V[prims[off:off + cnt]]
            assert np.array_equal(tb.min((0, 1)), nd["pmin"])
            assert np.array_equal(tb.max((0, 1)), nd["pmax"])
    assert seen_leaf.all()
    # leaf rule + SAH rule at every internal node
    span = {}

    def rng_of(i):
        if i in span:
            return span[i]
        nd = nodes[i]
        if nd["tag"] == 0:
            span[i] = (int(nd["a"]), int(nd["a"] + nd["b"]))
        else:
            a, b = rng_of(nd["a"]), rng_of(nd["b"])
            assert a[1] == b[0]                    # left range precedes right
            span[i] = (a[0], b[1])
        return span[i]

    for i in range(n):
        nd = nodes[i]
        lo, hi = rng_of(i)
        cnt = hi - lo
        if nd["tag"] == 0:
            continue
        assert cnt > max_leaf and depth[i] < max_depth
        L_, R_ = rng_of(nd["a"]), rng_of(nd["b"])
        lc = cen[prims[L_[0]:L_[1]]]
        rc = cen[prims[R_[0]:R_[1]]]
        # the partition predicate c[axis] < split separates the children
        assert any(lc[:, k].max() < rc[:, k].min() for k in range(3))
        # SAH: the chosen split's float64 cost is the minimum over candidates
        ext = (nd["pmax"] - nd["pmin"]).astype(np.float64)
        sa = lambda e: 2 * (e[0] * e[1] + e[0] * e[2] + e[1] * e[2])
        best = np.inf
        for k in range(3):
            if ext[k] <= 0:
                continue
            b_ = np.clip(((cen[prims[lo:hi], k] - nd["pmin"][k]) * (bins / ext[k])).astype(int), 0, bins - 1)
            for s_ in range(bins - 1):
                left = b_ <= s_
                nl = left.sum()
                if nl == 0 or nl == cnt:
                    continue
                vb = V[prims[lo:hi]]
                el = vb[left].max((0, 1)).astype(np.float64) - vb[left].min((0, 1))
                er = vb[~left].max((0, 1)).astype(np.float64) - vb[~left].min((0, 1))
                best = min(best, 1 + (nl * sa(el) + (cnt - nl) * sa(er)) / sa(ext))
        el = (nodes[nd["a"]]["pmax"] - nodes[nd["a"]]["pmin"]).astype(np.float64)
        er = (nodes[nd["b"]]["pmax"] - nodes[nd["b"]]["pmin"]).astype(np.float64)
        chosen = 1 + ((L_[1] - L_[0]) * sa(el) + (R_[1] - R_[0]) * sa(er)) / sa(ext)
        assert chosen <= best * (1 + 1e-4) + 1e-9, (i, chosen, best)


def test_bvh_invariants_and_sah_rule(orc):
    data = _sphere_scene()
    P, I = data.concatenated()
    nodes, prims, st = orc.build_bvh(P, I)
    assert st["root"] == len(nodes) - 1
    _check_bvh(nodes, prims, P, I)
    d2, _ = scenes.cornell_box()
    P2, I2 = d2.concatenated()
    n2, p2, _ = orc.build_bvh(P2, I2)
    _check_bvh(n2, p2, P2, I2)


def test_bvh_degenerate_inputs(orc):
    """All centroids equal: every candidate cost is NaN, so the reference falls
    back to axis 0 / split 0.0, the partition is one-sided and the node is a
    leaf; a single triangle and an empty soup are single leaves."""
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]] * 10, np.float32)
    I = np.arange(30, dtype=np.uint32)
    nodes, prims, st = orc.build_bvh(P, I)
    assert len(nodes) == 1 and nodes[0]["tag"] == 0 and nodes[0]["b"] == 10
    nodes, prims, st = orc.build_bvh(P[:3], I[:3])
    assert len(nodes) == 1 and nodes[0]["b"] == 1
    nodes, prims, st = orc.build_bvh(np.zeros((0, 3), np.float32), np.zeros(0, np.uint32))
    assert len(nodes) == 1 and nodes[0]["b"] == 0


# ---------------------------------------------------------------------------
# BSDFs (Material/*.cu): closed forms
# ---------------------------------------------------------------------------
def _hemi_grid(n=400):
    th = (np.arange(n) + 0.5) / n * (np.pi / 2)
    ph = (np.arange(2 * n) + 0.5) / (2 * n) * 2 * np.pi
    T, Pp = np.meshgrid(th, ph, indexing="ij")
    w = np.stack([np.sin(T) * np.cos(Pp), np.sin(T) * np.sin(Pp), np.cos(T)], -1)
    dA = np.sin(T) * (np.pi / 2 / n) * (2 * np.pi / (2 * n))
    return w.reshape(-1, 3), dA.reshape(-1)


def test_lambert_closed_form(orc):
    m = scenes.diffuse((0.2, 0.5, 0.8))
    rng = np.random.default_rng(0)
    for _ in range(20):
        wo = rng.normal(size=3)
        wi = rng.normal(size=3)
        wi[2] = abs(wi[2])
        wo, wi = wo / np.linalg.norm(wo), wi / np.linalg.norm(wi)
        f, pdf = orc.bsdf_eval_local(m, wo, wi)
        assert np.allclose(f, np.array([0.2, 0.5, 0.8]) / np.pi * abs(wi[2]), rtol=1e-5)
        assert math.isclose(pdf, wi[2] / np.pi, rel_tol=1e-5)
    u = np.array([0.3, 0.7], np.float32)
    wi, f, pdf = orc.bsdf_sample_local(m, [0, 0, 1], u)
    r, phi = math.sqrt(0.3), 2 * math.pi * 0.7                    # Malley
    assert np.allclose(wi, [r * math.cos(phi), r * math.sin(phi), math.sqrt(1 - 0.3)], atol=1e-6)
    assert math.isclose(pdf, wi[2] / math.pi, rel_tol=1e-5)


def test_phong_closed_form_and_sample_pdf_quirk(orc):
    kd, ks, n = (0.3, 0.3, 0.3), (0.5, 0.5, 0.5), 20.0
    m = scenes.phong(kd, ks, n)
    f, pdf = orc.bsdf_eval_local(m, [0, 0, 1], [0, 0, 1])
    expect = np.array(kd) / np.pi + np.array(ks) * (n + 2) / (2 * np.pi)
    assert np.allclose(f, expect, rtol=1e-5)
    pd, ps = 0.3 / 0.8, 0.5 / 0.8
    assert math.isclose(pdf, pd / np.pi + ps * (n + 1) / (2 * np.pi), rel_tol=1e-5)
    # the diffuse branch (u.x < pD) returns the cosine pdf alone, not the mixture
    wi, _, spdf = orc.bsdf_sample_local(m, [0, 0, 1], [0.1, 0.4])
    assert math.isclose(spdf, wi[2] / np.pi, rel_tol=1e-5)
    _, mix = orc.bsdf_eval_local(m, [0, 0, 1], wi)
    c = float(wi[2])
    assert math.isclose(mix, pd * c / np.pi + ps * (n + 1) / (2 * np.pi) * c ** n, rel_tol=1e-4)
    assert not math.isclose(mix, spdf, rel_tol=1e-3)


def test_principled_pdf_mass_and_lobe_probability_quirk(orc):
    """Integral of the Principled pdf over the hemisphere at normal incidence
    = pD + pS * (GGX mass with theta_h < 45 deg), with the reference's
    mis-normalised probabilities pD = 1/(1+s), pS = s/(pD+s), s = lerp(.1,1,m);
    at low roughness the GGX mass below 45 deg is ~1 (D cos is normalised)."""
    w, dA = _hemi_grid(300)
    for metal in (0.0, 1.0):
        m = scenes.principled((0.5, 0.5, 0.5), 0.2, metal)
        s = 0.1 * (1 - metal) + metal
        pD = 1 / (1 + s)
        pS = s / (pD + s)
        tot = 0.0
        wo = np.array([0, 0, 1], np.float32)
        pdfs = np.array([orc.bsdf_eval_local(m, wo, wi)[1] for wi in w[::7]])
        tot = float((pdfs * dA[::7]).sum() * 7)
        assert abs(tot - (pD + pS)) < 0.02 * (pD + pS), (metal, tot, pD + pS)


def test_principled_sample_pdf_is_evaluate_pdf(orc):
    m = scenes.principled((0.7, 0.4, 0.2), 0.5, 1.0)
    rng = np.random.default_rng(2)
    for _ in range(50):
        wo = rng.normal(size=3)
        wo[2] = abs(wo[2]) + 0.1
        wo /= np.linalg.norm(wo)
        u = rng.uniform(size=2).astype(np.float32)
        wi, f, pdf = orc.bsdf_sample_local(m, wo, u)
        _, epdf = orc.bsdf_eval_local(m, wo, wi)
        assert pdf == epdf


def test_schlick_fresnel_limits(orc):
    """Metal (f0 = 0.04 in the reference's inverted lerp at metallic 1) at
    normal incidence / mirror configuration: the specular term carries
    F(cd=1) = f0 = 0.04 for every channel regardless of base colour."""
    a = scenes.principled((0.9, 0.1, 0.1), 0.3, 1.0)
    b = scenes.principled((0.1, 0.9, 0.1), 0.3, 1.0)
    fa, _ = orc.bsdf_eval_local(a, [0, 0, 1], [0, 0, 1])
    fb, _ = orc.bsdf_eval_local(b, [0, 0, 1], [0, 0, 1])
    assert np.allclose(fa, fb, rtol=1e-6)            # no diffuse, F = 0.04 gray
    assert np.allclose(fa, fa[0], rtol=1e-6)


# ---------------------------------------------------------------------------
# Film (Core/Sensor.cu:19-43)
# ---------------------------------------------------------------------------
def test_film_splat_closed_form(orc):
    s = np.zeros((4, 4, 3), np.float32)
    w = np.zeros((4, 4), np.float32)
    orc.film_splat(s, w, [1, 2, 3], 1.5, 1.5)        # pixel centre: weight 1, alone
    assert w[1, 1] == 1 and w.sum() == 1 and np.array_equal(s[1, 1], [1, 2, 3])
    s[:] = 0
    w[:] = 0
    orc.film_splat(s, w, [1, 1, 1], 1.75, 1.5)
    assert math.isclose(w[1, 1], 0.75) and math.isclose(w[1, 2], 0.25)
    assert math.isclose(w.sum(), 1.0)
    s[:] = 0
    w[:] = 0
    orc.film_splat(s, w, [np.nan, 0, 0], 1.5, 1.5)   # NaN samples dropped
    assert w.sum() == 0


# ---------------------------------------------------------------------------
# Integrator (Core/Integrator.cu:142-227): closed-form scenes
# ---------------------------------------------------------------------------
def _cam(pos, target, fov, size, sensor=None):
    return scenes.make_camera(dict(pos=pos, target=target, up=(0, 0, -1) if pos[1] > 0 and
                                   pos[0] == target[0] and pos[2] == target[2] else (0, 1, 0),
                                   fov=fov), sensor)


def test_integrator_emitter_and_env_only(orc):
    em = scenes.emitter((3.0, 2.0, 1.0))
    data = cr.SceneData(meshes=[scenes.quad((-5, -5, -1), (10, 0, 0), (0, 10, 0), 0)],
                        materials=[em], envRadiance=(0.5, 0.5, 0.5))
    data.emissiveIndex = [0, 1]
    cam = _cam((0, 0, 3), (0, 0, 0), 20, (16, 16))
    s = orc.Scene(data)
    Lr, _, _ = s.render_samples(cam, (16, 16), cr.Configs(), 0)
    assert np.allclose(Lr, [3.0, 2.0, 1.0])          # depth-0 emitter hit
    cam2 = _cam((0, 0, 3), (0, 0, 10), 20, (16, 16))  # looking away: env miss
    Lr, _, _ = s.render_samples(cam2, (16, 16), cr.Configs(), 0)
    assert np.allclose(Lr, 0.5)


def test_integrator_diffuse_plane_under_env_kill_quirk(orc):
    """Diffuse plane (albedo rho) seen from above, constant environment E, no
    emitters: one bounce, f/pdf = rho, then the unconditional kill (survival
    1 - rrThreshold) before the escaping ray adds E: E[L] = 0.95 * rho * E."""
    rho, E = 0.5, 2.0
    data = cr.SceneData(meshes=[scenes.quad((-50, 0, 50), (100, 0, 0), (0, 0, -100), 0)],
                        materials=[scenes.diffuse((rho,) * 3)], envRadiance=(E,) * 3)
    cam = _cam((0, 5, 0), (0, 0, 0), 30, (64, 64))
    s = orc.Scene(data)
    tot = []
    for k in range(8):
        Lr, _, _ = s.render_samples(cam, (64, 64), cr.Configs(maxDepth=10, rrThreshold=0.05), k)
        tot.append(Lr[..., 0])
    x = np.concatenate([t.ravel() for t in tot])
    assert set(np.unique(np.round(x, 5))) <= {0.0, round(rho * E, 5)}
    se = np.sqrt(0.95 * 0.05) * rho * E / np.sqrt(len(x))
    assert abs(x.mean() - 0.95 * rho * E) < 4 * se


def _rect_irradiance_factor(a, b, h):
    """Point form factor from a differential area to a parallel a x b
    rectangle at height h, the point below one corner (Siegel & Howell)."""
    X, Y = a / h, b / h
    return (X / math.sqrt(1 + X * X) * math.atan(Y / math.sqrt(1 + X * X))
            + Y / math.sqrt(1 + Y * Y) * math.atan(X / math.sqrt(1 + Y * Y))) / (2 * math.pi)


@pytest.mark.parametrize("mis", [True, False])
def test_integrator_direct_light_double_count(orc, mis):
    """Diffuse plane under a parallel square emitter, environment black.  At a
    point below the emitter centre, direct light D = rho * Le * F.  The
    reference adds emitter radiance on *every* hit, so the continuation ray
    that reaches the light counts it again (after the 5 % kill):
    E[L] = D + 0.95 * D for both the MIS and non-MIS estimators."""
    rho, Le, half, h = 0.6, 5.0, 0.5, 1.0
    mats = [scenes.diffuse((rho,) * 3), scenes.emitter((Le,) * 3)]
    data = cr.SceneData(meshes=[scenes.quad((-50, 0, 50), (100, 0, 0), (0, 0, -100), 0),
                                scenes.quad((-half, h, -half), (2 * half, 0, 0), (0, 0, 2 * half), 1)],
                        materials=mats)
    data.emissiveIndex = [2, 3]
    F = 4 * _rect_irradiance_factor(half, half, h)
    D = rho * Le * F
    cam = _cam((0, 0.6, 0), (0, 0, 0), 1.0, (32, 32))   # looks down through the light's gap
    s = orc.Scene(data)
    cfg = cr.Configs(maxDepth=10, rrThreshold=0.05, useMIS=mis)
    vals = np.concatenate([s.render_samples(cam, (32, 32), cfg, k)[0][..., 0].ravel()
                           for k in range(24)])
    se = vals.std() / math.sqrt(len(vals))
    assert abs(vals.mean() - 1.95 * D) < 4 * se + 1e-3 * D, (vals.mean(), 1.95 * D, se)


# ---------------------------------------------------------------------------
# Spectral transport (an addition to the reference): the CIE table and XYZ sum
# of the reference's BroadSpectrum (Core/Spectrum.cpp:37-45, SpectrumTabData.h),
# and the reduction of flat spectra to the (reference-pinned) RGB estimator.
# ---------------------------------------------------------------------------
def _oracle_cmf(orc):
    out = np.zeros((95, 3), np.float32)
    orc.lib().orc_cie_cmf(out.ctypes.data)
    return out


def _oracle_xyz(orc, s):
    s = np.ascontiguousarray(s, np.float32)
    out = np.zeros(3, np.float32)
    orc.lib().orc_spectrum_to_xyz(s.ctypes.data, out.ctypes.data)
    return out


# Table spacing: the reference's generator sampled 361..830 nm at 95 points.
TABLE_LAMBDA = 361.0 + np.arange(95) * 469.0 / 94.0


def test_cie_table_is_the_standard_observer(orc):
    """The table against the CIE 1931 2-degree observer's published features:
    y-bar peaks at 555 nm with 1.0, x-bar at 599 nm with 1.0622, z-bar at 446 nm
    with 1.7830; integral of y-bar (1 nm sum) = 106.857, the reference's
    normalisation; equal-energy white has X = Y = Z."""
    c = _oracle_cmf(orc).astype(np.float64)
    lam = TABLE_LAMBDA
    assert abs(lam[np.argmax(c[:, 1])] - 555) < 3 and abs(c[:, 1].max() - 1.0) < 1e-3
    assert abs(lam[np.argmax(c[:, 0])] - 599) < 3 and abs(c[:, 0].max() - 1.0622) < 2e-3
    assert abs(lam[np.argmax(c[:, 2])] - 446) < 3 and abs(c[:, 2].max() - 1.7830) < 2e-3
    assert c.min() >= 0.0 and c[-1].max() < 1e-5 and c[0].max() < 1e-3
    tot = c.sum(0)
    assert np.all(np.abs(tot / tot[1] - 1.0) < 2e-3)
    assert abs(tot[1] * 469.0 / 94.0 - 106.857) < 0.05
    # the product's table is the same data
    assert np.array_equal(cr.cie_cmf(), c.astype(np.float32))


def test_spectrum_to_xyz_reference_normalisation_and_z_fix(orc):
    """XYZ = table sum * kXYZStep (470/95) * kCIENormalizationY (1/106.857).
    The table is spaced 469/94 nm, so a flat spectrum 1 gets
    Y = (470/95) / (469/94) = 0.99158 of the 1 nm integral's 1.  A spectrum
    at bin 17 (446 nm) alone has Z/X = z-bar/x-bar = 5.14 there; the
    reference's x-bar-for-Z bug would give 1."""
    one = np.ones(95, np.float32)
    xyz = cr.spectrum_to_xyz(one)
    assert abs(xyz[1] - (470 / 95) / (469 / 94)) < 1e-4
    assert np.allclose(xyz, xyz[1], rtol=2e-3)               # equal-energy white
    assert np.array_equal(xyz, _oracle_xyz(orc, one))
    d = np.zeros(95, np.float32)
    d[17] = 2.0
    xb = cr.spectrum_to_xyz(d)
    assert abs(xb[2] / xb[0] - 1.782902 / 0.346678) < 1e-4
    assert abs(xb[1] - 2.0 * 0.031038 * (470 / 95) / 106.8569131841719) < 1e-7
    rng = np.random.default_rng(4)
    a, b = rng.random(95).astype(np.float32), rng.random(95).astype(np.float32)
    lhs = cr.spectrum_to_xyz(2.0 * a + 3.0 * b)
    rhs = 2.0 * cr.spectrum_to_xyz(a) + 3.0 * cr.spectrum_to_xyz(b)
    assert np.allclose(lhs, rhs, rtol=1e-5)
    assert np.array_equal(cr.spectrum_to_xyz(a), _oracle_xyz(orc, a))


def test_flat_spectra_reduce_to_the_rgb_estimator(orc):
    """Spectral transport with constant spectra equal to gray RGB materials:
    the expected XYZ is the RGB estimator's gray value times the white point
    (X, 1, Z of a constant spectrum 1); per-pass image means agree within 3
    sigma (the two modes draw different random streams)."""
    data, camd = scenes.cornell_box(light_radiance=(12.0, 12.0, 12.0))
    grays = [0.73, 0.5, 0.3, 0.0]
    for i, g in enumerate(grays[:3]):
        data.materials[i] = scenes.diffuse((g, g, g))
    sp = np.zeros((4, 3, 95), np.float32)
    for i, g in enumerate(grays[:3]):
        sp[i, 0] = g
    sp[3, 2] = 12.0
    data.spectra = sp
    white = cr.spectrum_to_xyz(np.ones(95, np.float32))
    osc = orc.Scene(data)
    cam = scenes.make_camera(camd)
    size = (24, 24)
    rgb_cfg = cr.Configs(maxDepth=6, rrThreshold=0.05, useMIS=True, rrStartDepth=3)
    spec_cfg = cr.Configs(maxDepth=6, rrThreshold=0.05, useMIS=True, rrStartDepth=3,
                          spectral=True)
    gm, sm = [], []
    for k in range(48):
        Lg, _, _ = osc.render_samples(cam, size, rgb_cfg, k)
        Ls, _, _ = osc.render_samples(cam, size, spec_cfg, 10000 + k)
        gm.append(Lg[..., 1].mean())
        sm.append(Ls.reshape(-1, 3).mean(0))
    gm, sm = np.array(gm), np.array(sm)
    for ch in range(3):
        exp = gm * white[ch]
        se = np.sqrt(exp.var(ddof=1) / len(exp) + sm[:, ch].var(ddof=1) / len(sm))
        assert abs(exp.mean() - sm[:, ch].mean()) <= 3 * se, (ch, exp.mean(), sm[:, ch].mean(), se)
