// Exercises the C++ host API (include/craystal/CRaystal.h) the way the
// reference's host code uses its classes (Source/CRayli/CRayli.cpp:7-40).
// Without arguments: host-only checks (BVH build, material finalize, camera,
// errors).  With --gpu: also builds a Scene, renders 2 spp and traces rays.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "craystal/CRaystal.h"

using namespace CRay;

#define EXPECT(c)                                                      \
    do {                                                               \
        if (!(c)) {                                                    \
            std::fprintf(stderr, "FAILED %s at line %d\n", #c, __LINE__); \
            return 1;                                                  \
        }                                                              \
    } while (0)

static MeshData quad(Float3 c, Float3 a, Float3 b, Float3 n, uint32_t mat) {
    MeshData m;
    m.position = {c, c + a, c + a + b, c + b};
    m.normal = {n, n, n, n};
    m.index = {0, 1, 2, 0, 2, 3};
    m.materialID = mat;
    return m;
}

int main(int argc, char** argv) {
    bool gpu = argc > 1 && std::strcmp(argv[1], "--gpu") == 0;
    SceneData data;
    MaterialData white;
    white.diffuseRefl = {0.7f, 0.7f, 0.7f};
    white.finalize();
    MaterialData light;
    light.emission = {10.f, 10.f, 10.f};
    light.finalize();
    EXPECT(light.isEmissive() && !white.isEmissive() && white.isDiffuse());
    data.materials = {white, light};
    data.meshes.push_back(quad({-2, 0, 2}, {4, 0, 0}, {0, 0, -4}, {0, 1, 0}, 0));
    data.meshes.push_back(quad({-0.5f, 2, -0.5f}, {1, 0, 0}, {0, 0, 1}, {0, -1, 0}, 1));
    data.emissiveIndex = {2, 3};

    BVH bvh;
    bvh.build(data);
    crh_bvh_stats st = bvh.stats();
    EXPECT(st.prim_count == 4 && st.node_count == 1 && st.root == 0);
    std::vector<crh_bvh_node_ref> nodes;
    std::vector<uint32_t> prims;
    bvh.exportNodes(nodes, prims);
    EXPECT(nodes.size() == 1 && nodes[0].tag == 0 && nodes[0].b == 4);

    Camera cam;
    cam.setWorldPosition({0, 1, 4});
    cam.setTarget({0, 0.5f, 0});
    cam.calculateCameraData();
    EXPECT(std::fabs(cam.getData().w[2] + 0.992278f) < 1e-4f);

    bool threw = false;
    try {
        SceneData bad = data;
        bad.meshes[0].index[0] = 99;
        BVH b2;
        b2.build(bad);
    } catch (const RuntimeError& e) {
        threw = e.code() == CRH_ERR_INVALID_ARG;
    }
    EXPECT(threw);
    // RGB -> spectrum uplift: grays are flat, white's XYZ is the flat spectrum's
    float gray[3] = {0.25f, 0.25f, 0.25f}, sp[95], xyz[3], xyzFlat[3], flat[95];
    EXPECT(crh_rgb_to_spectrum(gray, sp) == CRH_OK);
    for (float v : sp) EXPECT(v == 0.25f);
    for (float& v : flat) v = 1.f;
    float one[3] = {1.f, 1.f, 1.f};
    EXPECT(crh_rgb_to_spectrum(one, sp) == CRH_OK);
    EXPECT(crh_spectrum_to_xyz(sp, xyz) == CRH_OK && crh_spectrum_to_xyz(flat, xyzFlat) == CRH_OK);
    for (int k = 0; k < 3; ++k) EXPECT(std::fabs(xyz[k] - xyzFlat[k]) < 1e-6f);
    if (!gpu) {
        std::printf("cpp host api ok\n");
        return 0;
    }
    // Spheres with their own materials and a 2 x 2 texture on the floor.
    SceneData sdata = data;
    MaterialData tex = white;
    tex.diffuseHandle = 0;
    tex.flags |= CRH_MATF_DIFFUSE_TEXTURE;
    sdata.materials.push_back(tex);
    sdata.meshes[0].texCrd = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
    sdata.meshes[0].materialID = 2;
    const float texel[12] = {0.9f, 0.1f, 0.1f, 0.1f, 0.9f, 0.1f,
                             0.1f, 0.1f, 0.9f, 0.9f, 0.9f, 0.9f};
    Image texImg(2, 2, 3);
    for (int i = 0; i < 4; ++i)
        for (int c = 0; c < 3; ++c) texImg.setPixel(i % 2, i / 2, c, texel[3 * i + c]);
    sdata.textureImages.push_back(std::move(texImg));
    sdata.spheres = {SphereData{{0.f, 0.5f, 0.f}, 0.4f}, SphereData{{1.f, 0.3f, 1.f}, 0.3f}};
    sdata.sphereMaterials = {0, 2};

    auto scene = std::make_shared<Scene>(std::move(data));
    auto sensor = std::make_shared<Sensor>(UInt2{32, 24});
    auto camRef = std::make_shared<Camera>(cam);
    camRef->setSensor(sensor);
    scene->setCamera(camRef);
    PathTraceIntegrator::Configs cfg;
    cfg.maxDepth = 4;
    PathTraceIntegrator integ(cfg);
    crh_stats rs = integ.dispatch(*scene, 2);
    auto img = sensor->createImage();
    EXPECT(img->getWidth() == 32 && img->getHeight() == 24 && img->getChannels() == 3);
    double mean = 0;
    for (const auto& layer : img->getRawData())
        for (float v : layer) mean += v;
    mean /= 3.0 * img->getArea();
    EXPECT(rs.rays > 0 && std::isfinite(mean) && mean > 0);
    std::vector<crh_ray> rays(2);
    rays[0] = crh_ray{{0, 1, 0}, 0.f, {0, -1, 0}, INFINITY};
    rays[1] = crh_ray{{0, 1, 0}, 0.f, {0, 1, 0}, INFINITY};
    auto hits = scene->intersect(rays);
    EXPECT(hits[0].prim < 2 && std::fabs(hits[0].t - 1.f) < 1e-6f);
    EXPECT(hits[1].prim >= 2 && std::fabs(hits[1].t - 1.f) < 1e-6f);
    auto occ = scene->intersectOcclusion(rays);
    EXPECT(occ[0] == 1 && occ[1] == 1);

    // Objects keep their context alive: destroy the context first, then use
    // and destroy a buffer made on it.
    crh_context* c2 = nullptr;
    crh_buffer* b2 = nullptr;
    EXPECT(crh_context_create(0, &c2) == CRH_OK);
    EXPECT(crh_buffer_create(c2, 64, &b2) == CRH_OK);
    EXPECT(crh_context_destroy(c2) == CRH_OK);
    uint32_t in[16], out[16] = {};
    for (int i = 0; i < 16; ++i) in[i] = 0x9E3779B9u * (i + 1);
    EXPECT(crh_buffer_upload(b2, in, 64) == CRH_OK);
    EXPECT(crh_buffer_download(b2, out, 64) == CRH_OK);
    EXPECT(std::memcmp(in, out, 64) == 0);
    EXPECT(crh_buffer_destroy(b2) == CRH_OK);

    auto sscene = std::make_shared<Scene>(std::move(sdata));
    std::vector<crh_ray> sr = {crh_ray{{0, 2, 0}, 0.f, {0, -1, 0}, INFINITY},
                               crh_ray{{0, 0.5f, 0}, 0.f, {1, 0, 0}, INFINITY}};
    auto sh = sscene->intersect(sr);
    EXPECT(sh[0].prim == (0u | CRH_HIT_SPHERE) && std::fabs(sh[0].t - 1.1f) < 1e-5f);
    EXPECT(sh[1].prim == (0u | CRH_HIT_SPHERE) && std::fabs(sh[1].t - 0.4f) < 1e-6f);  // from inside
    auto ssensor = std::make_shared<Sensor>(UInt2{32, 24});
    auto scam = std::make_shared<Camera>(cam);
    scam->setSensor(ssensor);
    sscene->setCamera(scam);
    PathTraceIntegrator::Configs scfg = cfg;
    scfg.spectral = true;
    crh_stats ss = PathTraceIntegrator(scfg).dispatch(*sscene, 2);
    auto simg = ssensor->createImage();
    double smean = 0;
    for (const auto& layer : simg->getRawData())
        for (float v : layer) smean += v;
    smean /= 3.0 * simg->getArea();
    EXPECT(ss.rays > 0 && std::isfinite(smean) && smean > 0 && std::fabs(smean - mean) > 1e-6);
    std::printf("cpp gpu api ok: mean %.5f rays %llu\n", mean, (unsigned long long)rs.rays);
    return 0;
}
