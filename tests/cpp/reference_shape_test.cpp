// Host code written the way the reference's own caller is (CRayli.cpp:7-42,
// which includes "CRaystal.h", uses namespace CRay and calls
// crayRenderSample), compiled against this library's headers: the types and
// calls it uses must keep their reference signatures -- Float3 / UInt2 with
// .x/.y/.z, Spectrum::fromRGB, Image::Ref Sensor::createImage() and
// Image::writeEXR.  Without arguments it builds the scene only; with an
// output path it also renders 4 spp and writes the EXR there (needs a GPU).
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <memory>

#include "craystal/CRaystal.h"

using namespace CRay;

void crayRenderSample(const PathTraceIntegrator::Ref& pIntegrator, Scene::Ref pScene, int spp,
                      const std::filesystem::path& outPath) {
    pIntegrator->dispatch(*pScene, spp);

    auto pImage = pScene->getCamera()->getSensor()->createImage();
    pImage->writeEXR(outPath.string());
}

static MeshData box(Float3 lo, Float3 hi, uint32_t mat) {
    MeshData m;
    for (int i = 0; i < 8; ++i)
        m.position.push_back(Float3(i & 1 ? hi.x : lo.x, i & 2 ? hi.y : lo.y, i & 4 ? hi.z : lo.z));
    const uint32_t f[36] = {0, 2, 1, 1, 2, 3, 4, 5, 6, 5, 7, 6, 0, 1, 4, 1, 5, 4,
                            2, 6, 3, 3, 6, 7, 0, 4, 2, 2, 4, 6, 1, 3, 5, 3, 7, 5};
    m.index.assign(f, f + 36);
    Float3 c = (lo + hi) * 0.5f;
    for (const Float3& p : m.position) m.normal.push_back(normalize(p - c));
    m.materialID = mat;
    return m;
}

int main(int argc, char const* argv[]) {
    Logger::init();
    Spectrum::initialize();

    SceneData data;
    MaterialData floor, lamp;
    floor.diffuseRefl = Spectrum::fromRGB(Float3(0.7f, 0.6f, 0.5f));
    floor.type = MaterialType::FullDiffuse;
    floor.finalize();
    lamp.emission = Spectrum(Float3(8.0f));
    lamp.finalize();
    data.materials = {floor, lamp};
    data.meshes.push_back(box(Float3(-2, -0.1f, -2), Float3(2, 0, 2), 0));
    data.meshes.push_back(box(Float3(-0.3f, 1.5f, -0.3f), Float3(0.3f, 1.6f, 0.3f), 1));
    for (uint32_t t = 12; t < 24; ++t) data.emissiveIndex.push_back(t);
    if (data.materials[0].diffuseRefl.r != 0.7f || data.meshes[1].position[7].y != 1.6f ||
        !data.materials[1].isEmissive()) {
        std::fprintf(stderr, "FAILED member access\n");
        return 1;
    }

    auto pCamera = std::make_shared<Camera>();
    pCamera->setFovY(45.f);
    pCamera->setWorldPosition(Float3(0, 1, 4));
    pCamera->setTarget(Float3(0, 0.5f, 0));
    pCamera->setUp(Float3(0, 1, 0));
    pCamera->calculateCameraData();
    UInt2 size(40u, 30u);
    if (argc < 2) {  // the film lives on the GPU: stop before it
        std::printf("reference-shaped host code compiles\n");
        return 0;
    }
    pCamera->setSensor(std::make_shared<Sensor>(size));
    if (pCamera->getSensor()->getSize().x != 40u || pCamera->getSensor()->getSize().y != 30u) {
        std::fprintf(stderr, "FAILED sensor size\n");
        return 1;
    }
    auto pScene = std::make_shared<Scene>(std::move(data));
    pScene->setCamera(pCamera);
    PathTraceIntegrator::Configs integratorConf;
    integratorConf.useMIS = true;
    auto pIntegrator = std::make_shared<PathTraceIntegrator>(integratorConf);
    std::filesystem::path outPath = argv[1];
    crayRenderSample(pIntegrator, pScene, 4, outPath);
    auto img = pScene->getCamera()->getSensor()->createImage();
    const float mid = img->getPixel(20, 15, 1);
    std::printf("rendered %s: pixel (20, 15) g = %.5f\n", outPath.string().c_str(), mid);
    return std::isfinite(mid) ? 0 : 1;
}
