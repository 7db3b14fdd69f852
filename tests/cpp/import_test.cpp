// Host-only check of include/craystal/Importer.h: imports scene.xml given on
// the command line and prints a summary line; writes a synthetic gradient
// image with writeEXR for byte comparison with the Python writer.
#include <cstdio>
#include <vector>

#include "craystal/Importer.h"

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    CRay::Importer im;
    CRay::CameraDesc cam;
    CRay::SceneData d = im.loadData(argv[1], cam);
    size_t tris = 0;
    for (const auto& m : d.meshes) tris += m.index.size() / 3;
    std::printf("meshes %zu materials %zu triangles %zu emissive %zu size %u %u fov %.4f\n",
                d.meshes.size(), d.materials.size(), tris, d.emissiveIndex.size(), cam.size[0],
                cam.size[1], cam.fovY);
    for (const auto& m : d.meshes) {
        std::printf("mesh mat %u verts %zu idx", m.materialID, m.position.size());
        for (uint32_t i : m.index) std::printf(" %u", i);
        std::printf(" pos");
        for (const auto& p : m.position) std::printf(" %.9g %.9g %.9g", p[0], p[1], p[2]);
        std::printf("\n");
    }
    for (uint32_t e : d.emissiveIndex) std::printf("em %u\n", e);
    const uint32_t w = 7, h = 5;
    std::vector<float> img(size_t(w) * h * 3);
    for (size_t i = 0; i < img.size(); ++i) img[i] = float(i) * 0.37f - 3.0f + 1e-6f * float(i * i);
    img[4] = 70000.f;   // overflows to inf in half
    img[5] = 1e-6f;     // half subnormal
    img[6] = -2.5e-8f;  // rounds to -0
    CRay::writeEXR(argv[2], w, h, img);
    return 0;
}
