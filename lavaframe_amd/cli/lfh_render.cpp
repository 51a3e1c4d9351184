// Headless render CLI with the reference's command-line surface for batch
// rendering (LavaFrame/Main.cpp:768-958): -s/--scene, -ms/--maxsamples,
// -ou/--outputpath, -ep/-exportname, -et/--exporttype, -db/--debug,
// -u/-nm/-w/-n (window / UI flags, accepted and ignored: always headless).
// It loads the scene with this build's loader, renders with HipRendererT
// through the C ABI until maxSamples passes are complete, and writes the
// linear HDR image.  Export formats: pfm (linear RGB float) and ppm (RGB8,
// clamped); the reference's png/jpg/tga/bmp/exr writers are display-side
// and outside this build.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../csrc/host_scene.h"
#include "lfh.h"
#include "lfh_renderer.hpp"

int main(int argc, char** argv) {
    std::string scene_file, out_path = "./", name = "", type = "pfm";
    int max_samples = -1;
    bool debug = false, device_bvh = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&](const char* what) -> std::string {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", what);
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "-s" || a == "--scene") scene_file = next("-s");
        else if (a == "-ms" || a == "--maxsamples") max_samples = std::atoi(next("-ms").c_str());
        else if (a == "-ou" || a == "--outputpath") out_path = next("-ou");
        else if (a == "-ep" || a == "-exportname") name = next("-ep");
        else if (a == "-et" || a == "--exporttype") type = next("-et");
        else if (a == "-db" || a == "--debug") debug = true;
        else if (a == "-dbvh" || a == "--device-bvh") device_bvh = true;
        else if (a == "-u" || a == "--noui" || a == "-nm" || a == "--nomove" || a == "-w" || a == "--nowindow" ||
                 a == "-n" || a == "--neutral") {
        } else if (a == "-h" || a == "--help") {
            std::printf("lfh_render -s scene.lfs -ms samples [-ou dir] [-ep name] [-et pfm|ppm] [-db] [-dbvh]\n"
                        "  -dbvh / --device-bvh: build the BVHs on the GPU instead of uploading the scene's\n");
            return 0;
        } else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    if (scene_file.empty()) {
        std::fprintf(stderr, "no scene given (-s)\n");
        return 2;
    }
    if (max_samples <= 0) max_samples = 64;
    lfh::Scene scene;
    std::string err;
    if (!lfh::LoadSceneFromFile(scene_file, &scene, scene.renderOptions, &err)) {
        std::fprintf(stderr, "could not load %s: %s\n", scene_file.c_str(), err.c_str());
        return 1;
    }
    scene.camera->isMoving = false;
    lfh::HipRendererT<lfh::Scene> renderer(&scene);
    renderer.SetDeviceBvh(device_bvh);
    try {
        renderer.Init();
        auto t0 = std::chrono::steady_clock::now();
        // Main loop order (Main.cpp:747-753): Update, then Render; export when
        // the sample counter reaches maxSamples + 1, i.e. after maxSamples passes.
        for (;;) {
            renderer.Update(0.f);
            if (renderer.GetSampleCount() == max_samples + 1) break;
            renderer.Render();
        }
        lfh_synchronize(renderer.context());
        double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (debug) std::printf("%d samples in %.3f s\n", max_samples, s);
        // export after maxSamples complete passes: the post-processed image of
        // accum / maxSamples (the sample counter already reads maxSamples + 1)
        const int w = scene.renderOptions.resolution.x, h = scene.renderOptions.resolution.y;
        lfh_postprocess pp = renderer.postprocess();
        const float inv = 1.0f / (float)max_samples;
        if (name.empty()) {
            size_t p = scene_file.find_last_of("/\\");
            name = scene_file.substr(p == std::string::npos ? 0 : p + 1);
            size_t d = name.find_last_of('.');
            if (d != std::string::npos) name = name.substr(0, d);
        }
        std::string path = out_path + (out_path.empty() || out_path.back() == '/' ? "" : "/") + name + "." + type;
        std::vector<float> hdr;
        std::vector<unsigned char> ldr;
        int rc;
        if (type == "ppm") {
            ldr.resize((size_t)w * h * 3);
            rc = lfh_get_output_ldr(renderer.context(), &pp, inv, ldr.data(), w, h);
        } else {
            hdr.resize((size_t)w * h * 3);
            rc = lfh_get_output_post(renderer.context(), &pp, inv, hdr.data(), w, h);
        }
        if (rc != LFH_OK) {
            std::fprintf(stderr, "export: %s\n", lfh_last_error(renderer.context()));
            return 1;
        }
        FILE* f = std::fopen(path.c_str(), "wb");
        if (!f) {
            std::fprintf(stderr, "cannot write %s\n", path.c_str());
            return 1;
        }
        if (type == "ppm") {
            // PPM rows are top-to-bottom
            std::fprintf(f, "P6\n%d %d\n255\n", w, h);
            for (int y = h - 1; y >= 0; --y) std::fwrite(&ldr[(size_t)y * w * 3], 1, (size_t)w * 3, f);
        } else {
            // PFM rows are bottom-to-top, like the GL readback
            std::fprintf(f, "PF\n%d %d\n-1.0\n", w, h);
            std::fwrite(hdr.data(), sizeof(float), hdr.size(), f);
        }
        std::fclose(f);
        std::printf("%s\n", path.c_str());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
