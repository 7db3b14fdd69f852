"""Closed-form pins of the oracle's Principled BRDF evaluate
(Material/PrincipledBRDF.cu:9-33,57-73,135-150, Material/Fresnel.cu:3-8,
built from (kd, ks, n) as Material/BSDF.cu:21-24 and :51-55 do).

The reference's evaluate is GGX D (alpha = r^2), Smith G with the Disney
remap alpha_g = (0.5 + r/2)^2 written as 1/(cos + sqrt(...)) per side (so it
already holds 1/(4 ci cv)) and then divided by 4 ci cv a second time, Schlick
Fresnel with the inverted lerp f0 = base (1 - m) + 0.04 m, and Disney diffuse
with fd90 = 0.5 + 2 r cd^2, times (1 - m).  BSDFBase::evaluate multiplies by
|cos theta_i| (Material/BSDF.h:15-58).

Each case below reduces that to a value worked by hand, so a dropped or
swapped term changes it; `test_principled_kats_catch_planted_mistakes`
compiles the oracle with each of five planted mistakes (ORC_MUTANT 1-5 in
oracle/oracle.cpp: the second division, the G remap, the f0 lerp, the 2 of
fd90, alpha in D) and checks every one fails at least one case.
"""
from __future__ import annotations

import ctypes as C
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from craystal_amd import scenes

ROOT = Path(__file__).resolve().parents[1]


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def _dir(theta_deg, phi_deg):
    t, p = math.radians(theta_deg), math.radians(phi_deg)
    return np.array([math.sin(t) * math.cos(p), math.sin(t) * math.sin(p), math.cos(t)])


def _roughness_of(n):
    # roughness = sqrt(2/(n+1)) in single precision (PrincipledBRDF.cu:54)
    return float(np.sqrt(np.float32(2.0) / (np.float32(n) + np.float32(1.0))))


def _cases():
    """(name, base, shininess, metallic, wo, wi, expected f*cos) with the
    expectation worked from a reduced closed form."""
    out = []
    # (1) r = 1 (shininess 1): alpha = 1 so D = 1/pi; alpha_g = 1 so each
    # Smith factor is 1/(cos + 1).  f*cos = ci * [diffuse] + F / (4 pi cv
    # (1 + ci)(1 + cv)) since D F G/(4 ci cv) * ci with G = 1/((1+ci)(1+cv)).
    for base, metal in (((0.7, 0.3, 0.1), 1.0), ((0.8, 0.5, 0.2), 0.5), ((0.6, 0.6, 0.6), 0.25)):
        for (to, po), (ti, pi_) in (((0, 0), (30, 90)), ((40, 10), (60, 200)),
                                     ((75, 0), (20, 180)), ((10, 45), (80, 300))):
            wo, wi = _dir(to, po), _dir(ti, pi_)
            cv, ci = wo[2], wi[2]
            wh = _unit(wo + wi)
            cd = float(wo @ wh)
            b = np.array(base)
            f0 = b * (1 - metal) + 0.04 * metal
            F = f0 + (1 - f0) * (1 - cd) ** 5
            spec_cos = F / (4 * math.pi * cv * (1 + ci) * (1 + cv))
            fd90 = 0.5 + 2 * 1.0 * cd * cd
            diff = b / math.pi * (1 + (fd90 - 1) * (1 - ci) ** 5) * \
                (1 + (fd90 - 1) * (1 - cv) ** 5) * (1 - metal)
            out.append((f"r1-{base}-{metal}-{to},{ti}", base, 1.0, metal, wo, wi,
                        diff * ci + spec_cos))
    # (2) normal incidence wo = wi = n: D = 1/(pi r^4), G = 1/4 (each factor
    # 1/(1 + 1)), F = f0, Disney factors 1:  f = base (1 - m)/pi + f0/(16 pi r^4).
    for r in (0.3, 0.7):
        n = 2.0 / (r * r) - 1.0
        rr = _roughness_of(n)
        for base, metal in (((0.9, 0.4, 0.2), 0.2), ((0.3, 0.6, 0.9), 0.0), ((0.5, 0.5, 0.5), 1.0)):
            b = np.array(base)
            f0 = b * (1 - metal) + 0.04 * metal
            f = b * (1 - metal) / math.pi + f0 / (16 * math.pi * rr ** 4)
            z = np.array([0.0, 0.0, 1.0])
            out.append((f"normal-r{r}-{base}-{metal}", base, n, metal, z, z, f))
    # (3) wi = n (ci = 1), grazing views: fl = 1 and fv = 1 + (fd90 - 1)(1 - cv)^5
    # with fd90 = 0.5 + 2 r cd^2; D from cos theta_h = (cv + 1)/|wo + n|; the
    # Smith factors in the textbook form G1 = 2c/(c + sqrt(ag^2 + (1 - ag^2) c^2))
    # (Walter et al. 2007) with ag = (0.5 + r/2)^2, so the reference's
    # product of 1/(c + sqrt(...)) factors is G1(ci) G1(cv) / (4 ci cv).
    for r in (0.25, 0.5, 0.9):
        n = 2.0 / (r * r) - 1.0
        rr = _roughness_of(n)
        for base, metal in (((0.8, 0.2, 0.4), 0.0), ((0.4, 0.7, 0.3), 0.6)):
            for tv in (70.0, 84.0, 88.0):
                wo, wi = _dir(tv, 30.0), np.array([0.0, 0.0, 1.0])
                cv, ci = wo[2], 1.0
                wh = _unit(wo + wi)
                ch, cd = wh[2], float(wo @ wh)
                a = rr * rr
                D = a * a / (math.pi * (1 + (a * a - 1) * ch * ch) ** 2)
                ag = (0.5 + rr / 2) ** 2

                def g1(c):
                    return 2 * c / (c + math.sqrt(ag * ag + (1 - ag * ag) * c * c))

                b = np.array(base)
                f0 = b * (1 - metal) + 0.04 * metal
                F = f0 + (1 - f0) * (1 - cd) ** 5
                spec = D * F * (g1(ci) * g1(cv) / (4 * ci * cv)) / (4 * ci * cv)
                fd90 = 0.5 + 2 * rr * cd * cd
                diff = b / math.pi * (1 + (fd90 - 1) * (1 - cv) ** 5) * (1 - metal)
                out.append((f"graze-r{r}-{base}-{metal}-{tv}", base, n, metal, wo, wi,
                            (diff + spec) * ci))
    return out


def _run_kats(eval_local):
    """Failures of every case: [(name, got, expected)]."""
    bad = []
    for name, base, n, metal, wo, wi, expect in _cases():
        m = scenes.MaterialData(diffuseRefl=tuple(base), specularRefl=(metal,) * 3,
                                shininess=n, type=scenes.MaterialType.Principled).finalize()
        f, _ = eval_local(m, wo.astype(np.float32), wi.astype(np.float32))
        if not np.allclose(f, expect, rtol=2e-4, atol=1e-7):
            bad.append((name, f, expect))
    return bad


def test_principled_evaluate_closed_forms(orc):
    bad = _run_kats(orc.bsdf_eval_local)
    assert not bad, bad[:4]


def _mutant_eval(path: Path):
    lib = C.CDLL(str(path))
    fn = lib.orc_bsdf_eval_local
    fn.restype = None
    fn.argtypes = [C.c_void_p] * 5

    def ev(mat, wo, wi):
        m = mat.to_c()
        wo = np.ascontiguousarray(wo, np.float32)
        wi = np.ascontiguousarray(wi, np.float32)
        f = np.zeros(3, np.float32)
        pdf = np.zeros(1, np.float32)
        fn(C.cast(C.byref(m), C.c_void_p), wo.ctypes.data, wi.ctypes.data, f.ctypes.data,
           pdf.ctypes.data)
        return f, float(pdf[0])
    return ev


@pytest.mark.parametrize("mutant", [1, 2, 3, 4, 5])
def test_principled_kats_catch_planted_mistakes(tmp_path, mutant):
    """A mistake planted in the oracle (on the CPU) fails at least one case."""
    so = tmp_path / f"liboracle_m{mutant}.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                    "-fopenmp", "-fPIC", "-shared", f"-DORC_MUTANT={mutant}", "-o", str(so),
                    str(ROOT / "oracle" / "oracle.cpp")], check=True)
    bad = _run_kats(_mutant_eval(so))
    assert bad, f"planted mistake {mutant} passed every closed form"
