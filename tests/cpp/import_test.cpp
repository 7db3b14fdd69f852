// Host-only check of include/craystal/Importer.h and Image.h: imports
// scene.xml given on the command line and prints a summary line; writes a
// synthetic gradient image with Image::writeEXR for byte comparison with the
// Python writer; with a third argument, loads that image file with
// Image::load and prints its size and every value as hex floats; with a
// fourth, writes the loaded image to it with writePNG (gamma corrected).
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
                d.meshes.size(), d.materials.size(), tris, d.emissiveIndex.size(), cam.size.x,
                cam.size.y, cam.fovY);
    for (const auto& m : d.meshes) {
        std::printf("mesh mat %u verts %zu idx", m.materialID, m.position.size());
        for (uint32_t i : m.index) std::printf(" %u", i);
        std::printf(" pos");
        for (const auto& p : m.position) std::printf(" %.9g %.9g %.9g", p.x, p.y, p.z);
        std::printf("\n");
    }
    for (uint32_t e : d.emissiveIndex) std::printf("em %u\n", e);
    const uint32_t w = 7, h = 5;
    std::vector<float> img(size_t(w) * h * 3);
    for (size_t i = 0; i < img.size(); ++i) img[i] = float(i) * 0.37f - 3.0f + 1e-6f * float(i * i);
    img[4] = 70000.f;   // overflows to inf in half
    img[5] = 1e-6f;     // half subnormal
    img[6] = -2.5e-8f;  // rounds to -0
    CRay::Image out(w, h, 3);
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c) out.setPixel(int(x), int(y), c, img[(y * w + x) * 3 + c]);
    out.writeEXR(argv[2]);
    for (const auto& m : d.materials)
        if (m.flags & uint32_t(CRay::MaterialFlags::hasDiffuseTexture))
            std::printf("texture %u %dx%dx%d\n", m.diffuseHandle,
                        d.textureImages[m.diffuseHandle].getWidth(),
                        d.textureImages[m.diffuseHandle].getHeight(),
                        d.textureImages[m.diffuseHandle].getChannels());
    if (argc > 3) {
        CRay::Image t = CRay::Image::load(argv[3]);
        std::printf("image %d %d %d\n", t.getWidth(), t.getHeight(), t.getChannels());
        for (int c = 0; c < t.getChannels(); ++c)
            for (int i = 0; i < t.getArea(); ++i) std::printf("%a\n", t.getPixel(i, c));
        if (argc > 4) t.writePNG(argv[4], true);
    }
    return 0;
}
