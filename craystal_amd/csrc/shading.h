// Device shading: hit reconstruction, shading frame, the three BRDFs and the
// light sampling / MIS estimator of the reference megakernel.
//
// Restates Scene/Scene.cu:31-58 (createIntersection), Core/Frame.cu:6-31,
// Material/BSDF.cu:6-103, Material/{Lambertian,Phong,Principled}BRDF.cu,
// Material/Fresnel.cu, Math/Sampling.cu and Core/Integrator.cu:13-140,
// including the reference's arithmetic quirks (double-precision
// subexpressions where the reference mixes double literals, mis-normalised
// Principled lobe probabilities, the Phong chosen-lobe sample pdf).
// Spectrum = RGB (Core/Spectrum.h:326).
#pragma once
#include "traverse.h"
#include "spec_layout.h"

namespace crh {

struct Spec {
    float r, g, b;
};
CRH_D Spec sp(float s) { return {s, s, s}; }
CRH_D Spec sp3(const float* p) { return {p[0], p[1], p[2]}; }
CRH_D Spec operator+(Spec a, Spec b) { return {a.r + b.r, a.g + b.g, a.b + b.b}; }
CRH_D Spec operator-(Spec a, Spec b) { return {a.r - b.r, a.g - b.g, a.b - b.b}; }
CRH_D Spec operator*(Spec a, Spec b) { return {a.r * b.r, a.g * b.g, a.b * b.b}; }
CRH_D Spec operator*(Spec a, float s) { return {a.r * s, a.g * s, a.b * s}; }
CRH_D Spec operator*(float s, Spec a) { return {s * a.r, s * a.g, s * a.b}; }
CRH_D Spec operator/(Spec a, float s) { return {a.r / s, a.g / s, a.b / s}; }
CRH_D float smaxv(Spec a) { return smax(smax(a.r, a.g), a.b); }
CRH_D float savg(Spec a) { return ((0.0f + a.r) + a.g + a.b) / 3.0f; }
CRH_D bool sbad(Spec a) {
    return isnan(a.r) || isnan(a.g) || isnan(a.b) || isinf(a.r) ||
           isinf(a.g) || isinf(a.b);
}

// --- Spectral transport -----------------------------------------------------
// Three wavelength bins (of the kSpecBins bins) per path ride in the three
// lanes of Spec; Lanes::tab points at the scene's spectral table
// (spec_layout.h), nullptr for RGB transport.
struct Lanes {
    const float* tab;
    uint32_t b0, b1, b2;
};
CRH_D Lanes rgbLanes() { return Lanes{nullptr, 0u, 0u, 0u}; }
// Bins h, h + 32, h + 64 (mod 95) for the hero bin h: each lane's bin is
// uniform over the 95 bins.
CRH_D Lanes heroLanes(const float* tab, float u) {
    uint32_t h = min(uint32_t(u * float(kSpecBins)), uint32_t(kSpecBins - 1));
    return Lanes{tab, h, (h + 32u) % kSpecBins, (h + 64u) % kSpecBins};
}
CRH_D Spec laneSpec(const float* p, const Lanes& w) { return {p[w.b0], p[w.b1], p[w.b2]}; }
CRH_D const float* specMat(const Lanes& w, uint32_t mid) {
    return w.tab + kSpecMat0 + size_t(mid) * kSpecMatStride;
}
// XYZ of the three lanes' radiance (the table folds in 95/3, 5 nm and the
// normalisation that maps a constant spectrum 1 to Y = 1).
CRH_D Spec laneXYZ(const Spec& L, const Lanes& w) {
    const float4* c = reinterpret_cast<const float4*>(w.tab + kSpecCmf);
    float4 a = c[w.b0], b = c[w.b1], d = c[w.b2];
    return {(L.r * a.x + L.g * b.x) + L.b * d.x, (L.r * a.y + L.g * b.y) + L.b * d.y,
            (L.r * a.z + L.g * b.z) + L.b * d.z};
}

// --- Triangle data (Scene/Triangle.cu:9-37, 78-92) -----------------------
struct TriData {
    F3 p[3], n[3];
    F2 uv[3];
};
CRH_D TriData getTri(const SceneView& s, uint32_t id) {
    TriData t;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        uint32_t vi = s.idx[3 * id + k];
        t.p[k] = ld3(s.pos + 3 * size_t(vi));
        t.n[k] = ld3(s.nrm + 3 * size_t(vi));
        t.uv[k] = F2{s.uv[2 * size_t(vi)], s.uv[2 * size_t(vi) + 1]};
    }
    return t;
}
CRH_D F3 faceNormal(const TriData& t) {
    return normalize(cross(t.p[1] - t.p[0], t.p[2] - t.p[0]));
}
CRH_D F3 interp3(const F3* v, F3 w) {
    return v[0] * w.x + v[1] * w.y + v[2] * w.z;
}
CRH_D float triArea(const TriData& t) {
    return length(cross(t.p[0] - t.p[1], t.p[2] - t.p[1])) / 2.f;
}

// --- Frame (Core/Frame.cu:6-31) -------------------------------------------
struct Frame {
    F3 T, B, N;
};
CRH_D Frame makeFrame(F3 n) {
    F3 t = mk3(0, 1, 0);
    if (1.0 - double(absDot(n, t)) < double(kEps)) t = mk3(1, 0, 0);
    F3 bt = normalize(cross(n, t));
    Frame f;
    f.N = n;
    f.T = normalize(cross(bt, n));
    f.B = cross(f.N, f.T);
    return f;
}
CRH_D F3 toLocal(const Frame& f, F3 w) {
    return {f.T.x * w.x + f.T.y * w.y + f.T.z * w.z,
            f.B.x * w.x + f.B.y * w.y + f.B.z * w.z,
            f.N.x * w.x + f.N.y * w.y + f.N.z * w.z};
}
CRH_D F3 toWorld(const Frame& f, F3 l) {
    return {f.T.x * l.x + f.B.x * l.y + f.N.x * l.z,
            f.T.y * l.x + f.B.y * l.y + f.N.y * l.z,
            f.T.z * l.x + f.B.z * l.y + f.N.z * l.z};
}

// --- Intersection (Scene/Scene.cu:31-58, Core/Intersection.cu:4-10) -------
struct Isect {
    F3 pos, faceN, view;
    Frame frame;
    bool front;
};
CRH_D F3 oriented(const Isect& it) { return it.front ? it.faceN : -it.faceN; }

CRH_D Isect makeIsect(const SceneView& s, const Ray& ray, const Hit& h) {
    Isect it;
    it.pos = ray.o + h.t * ray.d;
    it.view = -normalize(ray.d);
    F3 sn;
    if (h.prim & kSphereBit) {  // HitType::Sphere: both normals radial
        const float* c = s.spheres + 4 * size_t(h.prim & ~kSphereBit);
        it.faceN = normalize(it.pos - mk3(c[0], c[1], c[2]));
        sn = it.faceN;
    } else {
        F3 w = mk3(1.f - h.u - h.v, h.u, h.v);
        TriData tri = getTri(s, h.prim);
        it.faceN = faceNormal(tri);
        sn = normalize(interp3(tri.n, w));
    }
    it.front = dot(it.view, it.faceN) > 0.0f;
    it.frame = makeFrame(it.front ? sn : -sn);
    return it;
}

// Material of a hit primitive: Core/Integrator.cu:172-178 reads the mesh
// descriptor of primitiveIndex whatever the hit type, so a sphere takes the
// material of the triangle with its index unless the scene gives spheres
// their own materials.
CRH_D uint32_t matIdOf(const SceneView& s, uint32_t prim) {
    if (prim & kSphereBit) {
        const uint32_t i = prim & ~kSphereBit;
        return s.sphereMat ? s.sphereMat[i] : s.primMat[i];
    }
    return s.primMat[prim];
}
// evalMIS's test that the BSDF ray found the sampled light
// (Core/Integrator.cu:108) compares primitive indices only; with the
// reference's sphere materials a sphere of the same index matches too.
CRH_D bool hitsLight(const SceneView& s, uint32_t prim, uint32_t eid) {
    return s.sphereMat ? prim == eid : (prim & ~kSphereBit) == eid;
}

// The hit's texture coordinate (Scene/Scene.cu:36-55): a triangle's
// interpolated vertex coordinates; a sphere's (u, v) parameterisation of the
// unit normal at the hit point (Scene/Sphere.cu:30-37).
CRH_D F2 hitUV(const SceneView& s, const Hit& h, const Isect& it) {
    if (h.prim & kSphereBit) {
        const F3 n = it.faceN;
        float theta = acosf(-n.y);
        float phi = atan2f(-n.z, n.x) + kPi;
        return F2{phi * (0.5f * kInvPi), theta * kInvPi};
    }
    F3 w = mk3(1.f - h.u - h.v, h.u, h.v);
    const uint32_t* vi = s.idx + 3 * size_t(h.prim);
    F2 t0{s.uv[2 * size_t(vi[0])], s.uv[2 * size_t(vi[0]) + 1]};
    F2 t1{s.uv[2 * size_t(vi[1])], s.uv[2 * size_t(vi[1]) + 1]};
    F2 t2{s.uv[2 * size_t(vi[2])], s.uv[2 * size_t(vi[2]) + 1]};
    return F2{t0.x * w.x + t1.x * w.y + t2.x * w.z, t0.y * w.x + t1.y * w.y + t2.y * w.z};
}

// --- Textures (Core/Texture.cu:6-35) ----------------------------------------
// float -> uint32 as the GPU converts: NaN and negatives 0, saturating.
CRH_D uint32_t satU32(float f) {
    return f >= 4294967296.0f ? 0xFFFFFFFFu : (f > 0.0f ? uint32_t(f) : 0u);
}
CRH_D float4 lerp4(float4 a, float4 b, float t) {
    const float r = 1.0f - t;
    return make_float4(a.x * r + b.x * t, a.y * r + b.y * t, a.z * r + b.z * t,
                       a.w * r + b.w * t);
}
// TextureView::sample: bilinear over texels clamped to the image.
CRH_D float4 sampleTexture(const SceneView& s, uint32_t tex, F2 uv) {
    const uint4 ti = reinterpret_cast<const uint4*>(s.texInfo)[tex];
    const float4* px = reinterpret_cast<const float4*>(s.texels) + ti.x;
    const uint32_t w = ti.y, h = ti.z;
    const float fxp = uv.x * float(w - 1u);
    const float fyp = uv.y * float(h - 1u);
    const uint32_t x0 = satU32(floorf(fxp)), y0 = satU32(floorf(fyp));
    const uint32_t x1 = min(x0 + 1u, w - 1u), y1 = min(y0 + 1u, h - 1u);
    const float fx = fxp - float(x0), fy = fyp - float(y0);
    const uint32_t cx0 = min(x0, w - 1u), cy0 = min(y0, h - 1u);
    const uint32_t cx1 = min(x1, w - 1u), cy1 = min(y1, h - 1u);
    float4 c00 = px[size_t(cy0) * w + cx0], c10 = px[size_t(cy0) * w + cx1];
    float4 c01 = px[size_t(cy1) * w + cx0], c11 = px[size_t(cy1) * w + cx1];
    return lerp4(lerp4(c00, c10, fx), lerp4(c01, c11, fx), fy);
}

// --- Sampling warps (Math/Sampling.cu) -------------------------------------
CRH_D F2 sampleBary(F2 u) {
    bool flip = (u.x + u.y) > 1.0f;
    return {flip ? 1.0f - u.x : u.x, flip ? 1.0f - u.y : u.y};
}
CRH_D F3 cosHemi(F2 u) {
    float r = sqrtf(u.x);
    float phi = 2.0f * kPi * u.y;
    float x = r * cosf(phi);
    float y = r * sinf(phi);
    float z = sqrtf(smax(0.0f, 1.0f - x * x - y * y));
    return {x, y, z};
}
CRH_D float cosHemiPdf(F3 d) { return smax(0.0f, d.z) * kInvPi; }
CRH_D float powerHeur(float f, float g) { return (f * f) / (f * f + g * g); }
CRH_D float acosSafe(float v) {
    const float lo = kEps, hi = float(1.0 - double(kEps));
    float c = v < lo ? lo : (hi < v ? hi : v);
    return acosf(c);
}

// --- BSDF (Material/BSDF.cu:6-103) -----------------------------------------
struct BSDF {
    Frame frame;
    Spec kd, ks;
    float n;
    float pD, pS;   // lobe probabilities (Phong, Principled)
    float rough, metal;
    uint32_t type;
};

// avgKd / avgKs: the reference's Spectrum::averageValue of the reflectances
// (over the three RGB channels, or over all 95 bins in spectral transport).
CRH_D BSDF makeBSDFWith(const GMaterial& m, const Frame& f, Spec kd, Spec ks, float avgKd,
                        float avgKs) {
    BSDF b;
    b.frame = f;
    b.type = m.type;
    b.kd = kd;
    b.ks = ks;
    b.n = m.shininess;
    b.pD = b.pS = 0.f;
    b.rough = b.metal = 0.f;
    if (m.type == 2u) {  // Phong (PhongBRDF.cu:10-19)
        float d = smax(avgKd, 0.0f), s = smax(avgKs, 0.0f);
        float sum = smax(d + s, kEps);
        b.pD = d / sum;
        b.pS = s / sum;
    } else if (m.type == 1u) {  // Principled (PrincipledBRDF.cu:47-58)
        float r = sqrtf(2.0f / (b.n + 1.0f));
        float mt = avgKs;
        b.rough = fminf(fmaxf(r, 0.0f), 1.0f);
        b.metal = fminf(fmaxf(mt, 0.0f), 1.0f);
        float spr = 0.1f * (1.0f - b.metal) + 1.0f * b.metal;
        float pd = 1.0f / (1.0f + spr);
        b.pD = pd;
        b.pS = spr / (pd + spr);
    }
    return b;
}

CRH_D BSDF makeBSDF(const GMaterial& m, const Frame& f) {
    Spec kd = sp3(m.kd), ks = sp3(m.ks);
    return makeBSDFWith(m, f, kd, ks, savg(kd), savg(ks));
}

// Smits's RGB -> spectrum decomposition (crh_rgb_to_spectrum) of one
// texel, evaluated at the path's three bins; avg = its average over all
// bins.  The basis rows and their averages live in the spectral table.
CRH_D Spec upliftLanes(const float* tab, const Lanes& w, float r, float g, float b, float& avg) {
    r = smax(r, 0.0f);
    g = smax(g, 0.0f);
    b = smax(b, 0.0f);
    float white, w1, w2;
    uint32_t b1, b2;
    if (r <= g && r <= b) {
        white = r;
        b1 = 0u;
        if (g <= b) { w1 = g - r; b2 = 5u; w2 = b - g; }
        else        { w1 = b - r; b2 = 4u; w2 = g - b; }
    } else if (g <= r && g <= b) {
        white = g;
        b1 = 1u;
        if (r <= b) { w1 = r - g; b2 = 5u; w2 = b - r; }
        else        { w1 = b - g; b2 = 3u; w2 = r - b; }
    } else {
        white = b;
        b1 = 2u;
        if (r <= g) { w1 = r - b; b2 = 4u; w2 = g - r; }
        else        { w1 = g - b; b2 = 3u; w2 = r - g; }
    }
    const float* p1 = tab + kSpecBasis + b1 * kSpecBins;
    const float* p2 = tab + kSpecBasis + b2 * kSpecBins;
    avg = white + w1 * tab[kSpecBasisAvg + b1] + w2 * tab[kSpecBasisAvg + b2];
    return {white + w1 * p1[w.b0] + w2 * p2[w.b0], white + w1 * p1[w.b1] + w2 * p2[w.b1],
            white + w1 * p1[w.b2] + w2 * p2[w.b2]};
}

// The BSDF at a hit (getBSDF, Material/BSDF.cu:6-32): the material's
// diffuse reflectance or its texture at the hit's coordinate, under the
// path's lanes.
CRH_D BSDF makeBSDF(const SceneView& s, const GMaterial& m, uint32_t mid, const Isect& it,
                    const Hit& h, const Lanes& w) {
    if (!(m.flags & 4u)) {
        if (!w.tab) return makeBSDF(m, it.frame);
        const float* sm = specMat(w, mid);
        return makeBSDFWith(m, it.frame, laneSpec(sm + kSpecKd, w), laneSpec(sm + kSpecKs, w),
                            sm[kSpecAvgKd], sm[kSpecAvgKs]);
    }
    const float4 t = sampleTexture(s, m.tex, hitUV(s, h, it));
    if (!w.tab) {
        Spec kd{t.x, t.y, t.z}, ks = sp3(m.ks);
        return makeBSDFWith(m, it.frame, kd, ks, savg(kd), savg(ks));
    }
    const float* sm = specMat(w, mid);
    float avgKd;
    Spec kd = upliftLanes(w.tab, w, t.x, t.y, t.z, avgKd);
    return makeBSDFWith(m, it.frame, kd, laneSpec(sm + kSpecKs, w), avgKd, sm[kSpecAvgKs]);
}
CRH_D Spec emissionOf(const GMaterial& m, uint32_t mid, const Lanes& w) {
    return w.tab ? laneSpec(specMat(w, mid) + kSpecEm, w) : sp3(m.emission);
}

// Phong (PhongBRDF.cu:21-68)
CRH_D Spec phongEval(const BSDF& b, F3 wo, F3 wi) {
    F3 r = reflect(-wo, mk3(0, 0, 1));
    return b.kd * kInvPi +
           b.ks * (b.n + 2) * kInv2Pi * powf(smax(dot(r, wi), 0.0f), b.n);
}
CRH_D float phongPdf(const BSDF& b, F3 wo, F3 wi) {
    float dp = cosHemiPdf(wi);
    F3 r = reflect(-wo, mk3(0, 0, 1));
    float ca = smax(dot(r, wi), 0.0f);
    float spdf = (b.n + 1) * kInv2Pi * powf(ca, b.n);
    return dp * b.pD + spdf * b.pS;
}
CRH_D F3 phongSample(const BSDF& b, F3 wo, F2 u, float& pdf) {
    F3 wi = mk3(0, 0, 0);
    pdf = 0.0f;
    if (b.pD > 0.0f && u.x < b.pD) {
        F2 u0{u.x / b.pD, u.y};
        wi = cosHemi(u0);
        pdf = cosHemiPdf(wi);
    } else if (b.pS > 0.0f) {
        float ux = (u.x - b.pD) / b.pS;
        float alpha = acosSafe(powf(ux, 1 / (b.n + 1)));
        float th = acosSafe(wo.z) + alpha;
        const float hiPi = float(double(kPi) / 2.0);
        th = th < 0.0f ? 0.0f : (hiPi < th ? hiPi : th);
        float phi = k2Pi * u.y;
        float st = sinf(th);
        wi = mk3(st * cosf(phi), st * sinf(phi),
                 float(sqrt(1.0 - double(st * st))));
        pdf = (b.n + 1) * kInv2Pi * powf(cosf(alpha), b.n);
    }
    return wi;
}

// Principled (PrincipledBRDF.cu:9-151)
CRH_D float ggxD(float r, float ch) {
    float a = r * r;
    float a2 = a * a;
    float c2 = ch * ch;
    float den = float(1.0 + (double(a2) - 1.0) * double(c2));
    return a2 * kInvPi / (den * den);
}
CRH_D float smithG1(float aG, float cv) {
    float a = aG * aG;
    float b = cv * cv;
    return 1 / (cv + sqrtf(a + b - a * b));
}
CRH_D float smithG(float r, float ci, float cv) {
    float a = float(0.5 + double(r) * 0.5);
    a *= a;
    return smithG1(a, ci) * smithG1(a, cv);
}
// The reference's pow(1.0 - c, 5.0) in double (PrincipledBRDF.cu:138-139)
// as ((x x)^2) x: within 2 double ulps of pow's value, far below the float
// rounding the result goes through next, at a few multiplications instead
// of a double-precision pow.
CRH_D double pow5d(double x) {
    const double x2 = x * x;
    return x2 * x2 * x;
}
CRH_D Spec prEval(const BSDF& b, F3 wo, F3 wi) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return sp(0.f);
    F3 wh = normalize(wo + wi);
    float cv = wo.z, ci = wi.z, ch = wh.z;
    float cd = dot(wo, wh);
    // evaluateDiffuse (:135-141), double subexpressions as in the reference
    float fd90 = float(0.5 + 2.0 * double(b.rough) * double(cd) * double(cd));
    float fl = float(1.0 + (double(fd90) - 1.0) * pow5d(1.0 - double(ci)));
    float fv = float(1.0 + (double(fd90) - 1.0) * pow5d(1.0 - double(cv)));
    Spec diff = b.kd * kInvPi * fv * fl * (1.0f - b.metal);
    // evaluateSpecular (:143-150)
    float D = ggxD(b.rough, ch);
    Spec f0 = b.kd * (1.0f - b.metal) + sp(0.04f) * b.metal;
    float v0 = 1.f - cd;
    float v1 = v0 * v0;
    Spec F = f0 + (sp(1.f) - f0) * v1 * v1 * v0;
    float G = smithG(b.rough, ci, cv);
    float den = float(4.0 * double(ci) * double(cv));
    Spec spec = D * F * G / den;
    return diff + spec;
}
CRH_D float prPdf(const BSDF& b, F3 wo, F3 wi) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
    F3 wh = normalize(wo + wi);
    float ch = wh.z;
    float cd = dot(wo, wh);
    if (cd <= 0.0f) return 0.0f;
    float dp = cosHemiPdf(wi);
    float spdf = 0.0f;
    if (ch > 0.0f) spdf = ggxD(b.rough, ch) * ch / (4.0f * cd);
    return dp * b.pD + spdf * b.pS;
}
CRH_D F3 prSample(const BSDF& b, F3 wo, F2 u, float& pdf) {
    if (wo.z <= 0.0f) {
        pdf = 0.0f;
        return mk3(0, 0, 0);
    }
    float u1 = u.x, u2 = u.y;
    F3 wi;
    if (u1 < b.pS) {
        u1 /= b.pS;
        float alpha = b.rough * b.rough;
        float a2 = alpha * alpha;
        float phi = u2 * k2Pi;
        float tan2 = a2 * u1 / (1 - u1);
        float ct = 1 / sqrtf(1 + tan2);
        float r = sqrtf(smax(1 - ct * ct, 0.0f));
        F3 wh = mk3(cosf(phi) * r, sinf(phi) * r, ct);
        wi = reflect(-wo, wh);
    } else {
        u1 = (u1 - b.pS) / (1.0f - b.pS);
        wi = cosHemi(F2{u1, u2});
    }
    pdf = prPdf(b, wo, wi);
    return wi;
}

// BSDFBase::evaluate multiplies by |cos| (BSDF.h:23-27).
CRH_D Spec evalLocal(const BSDF& b, F3 wo, F3 wi) {
    Spec f;
    if (b.type == 0u) f = b.kd * kInvPi;
    else if (b.type == 1u) f = prEval(b, wo, wi);
    else f = phongEval(b, wo, wi);
    return f * fabsf(wi.z);
}
CRH_D float pdfLocal(const BSDF& b, F3 wo, F3 wi) {
    if (b.type == 0u) return cosHemiPdf(wi);
    if (b.type == 1u) return prPdf(b, wo, wi);
    return phongPdf(b, wo, wi);
}
CRH_D F3 sampleLocal(const BSDF& b, F3 wo, F2 u, float& pdf) {
    if (b.type == 0u) {
        F3 wi = cosHemi(u);
        pdf = cosHemiPdf(wi);
        return wi;
    }
    if (b.type == 1u) return prSample(b, wo, u, pdf);
    return phongSample(b, wo, u, pdf);
}
CRH_D Spec bsdfEval(const BSDF& b, F3 wo, F3 wi) {
    return evalLocal(b, toLocal(b.frame, wo), toLocal(b.frame, wi));
}
CRH_D float bsdfPdf(const BSDF& b, F3 wo, F3 wi) {
    return pdfLocal(b, toLocal(b.frame, wo), toLocal(b.frame, wi));
}
CRH_D Spec bsdfSampleEval(const BSDF& b, Sampler& smp, F3 wo, F3& wi,
                          float& pdf) {
    F3 woL = toLocal(b.frame, wo);
    pdf = 0.0f;
    F2 u = smp.next2D();
    F3 wiL = sampleLocal(b, woL, u, pdf);
    Spec f = evalLocal(b, woL, wiL);
    wi = toWorld(b.frame, wiL);
    return f;
}

}  // namespace crh
