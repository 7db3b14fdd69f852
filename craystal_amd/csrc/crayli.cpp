// Command line renderer with the reference CLI's options (CRayli.cpp:16-42):
//   crayli modelXML [--spp N] [--no-mis] [-o out.exr]
// Imports scene.xml + scene.obj, renders with the configuration the
// reference's kernel runs with (max depth 10, Russian roulette threshold 0.05
// after depth 5; MIS unless --no-mis) and writes a half-float B, G, R EXR.
#include <cstdio>
#include <cstring>
#include <string>

#include "craystal/Importer.h"

int main(int argc, char** argv) {
    std::string model, out = "craystal.exr";
    int spp = 16;
    bool noMIS = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--spp" && i + 1 < argc) {
            spp = std::stoi(argv[++i]);
        } else if (a == "--no-mis") {
            noMIS = true;
        } else if ((a == "-o" || a == "--output") && i + 1 < argc) {
            out = argv[++i];
        } else if (a == "-h" || a == "--help") {
            std::printf("usage: crayli modelXML [--spp N] [--no-mis] [-o out.exr]\n");
            return 0;
        } else if (model.empty()) {
            model = a;
        } else {
            std::fprintf(stderr, "unexpected argument %s\n", a.c_str());
            return 2;
        }
    }
    if (model.empty()) {
        std::fprintf(stderr, "usage: crayli modelXML [--spp N] [--no-mis] [-o out.exr]\n");
        return 2;
    }
    try {
        CRay::Importer importer;
        auto scene = importer.import(model);
        auto cfg = CRay::PathTraceIntegrator::Configs::reference();
        cfg.useMIS = !noMIS;
        CRay::PathTraceIntegrator integrator(cfg);
        crh_stats st = integrator.dispatch(*scene, spp);
        // crayRenderSample (CRayli.cpp:7-14): the sensor's image as EXR
        scene->getCamera()->getSensor()->createImage()->writeEXR(out);
        std::printf("wrote %s: %d spp, %.1f ms, %llu rays\n", out.c_str(), spp, st.kernel_ms,
                    static_cast<unsigned long long>(st.rays));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
