// A C++ host of the library in the order the reference core would call a
// hip backend plugin's hooks: TEST INFRASTRUCTURE. It dlopens the library
// named on the command line, as a plugin's shared object would be loaded,
// and draws one still frame of a scene file:
//
//   openWorkers      -> kdn_open
//   buildWorld       -> kdn_scene_load, kdn_upload_scene, kdn_scene_region,
//                       kdn_build_octree  (core/render.cpp:187)
//   drawRound        -> kdn_render_pass with estimates (the whole-pass hook
//                       of a round; core/frame.cpp:453-467 runs each round)
//                    -> kdn_read_estimates (the frame's Estimate per pixel,
//                       core/stopping.h:57-100)
//   deliverFrame     -> kdn_image_write_ppm, kdn_image_write_pfm
//                       (sink/file, shared/image.h:104,124)
//   closeWorkers     -> kdn_scene_free, kdn_close
//
// usage: render_host LIBRARY SCENE.sdf OUT_STEM WIDTH HEIGHT SPP DEPTH
// Prints the frame's mean and the bytes it wrote; exits nonzero on any error
// code with the library's message.
#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "kodanuki.h"

namespace {

template <class F>
F symbol(void* lib, const char* name) {
    void* p = dlsym(lib, name);
    if (!p) {
        std::fprintf(stderr, "render_host: %s is not exported\n", name);
        std::exit(2);
    }
    return reinterpret_cast<F>(p);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 8) {
        std::fprintf(stderr, "usage: render_host LIBRARY SCENE.sdf OUT_STEM WIDTH HEIGHT SPP DEPTH\n");
        return 2;
    }
    void* lib = dlopen(argv[1], RTLD_NOW | RTLD_LOCAL);
    if (!lib) {
        std::fprintf(stderr, "render_host: %s\n", dlerror());
        return 2;
    }
    const auto scene_load = symbol<decltype(&kdn_scene_load)>(lib, "kdn_scene_load");
    const auto scene_free = symbol<decltype(&kdn_scene_free)>(lib, "kdn_scene_free");
    const auto scene_settings = symbol<decltype(&kdn_scene_settings)>(lib, "kdn_scene_settings");
    const auto scene_region = symbol<decltype(&kdn_scene_region)>(lib, "kdn_scene_region");
    const auto open = symbol<decltype(&kdn_open)>(lib, "kdn_open");
    const auto close = symbol<decltype(&kdn_close)>(lib, "kdn_close");
    const auto last_error = symbol<decltype(&kdn_last_error)>(lib, "kdn_last_error");
    const auto global_error = symbol<decltype(&kdn_global_error)>(lib, "kdn_global_error");
    const auto upload = symbol<decltype(&kdn_upload_scene)>(lib, "kdn_upload_scene");
    const auto build = symbol<decltype(&kdn_build_octree)>(lib, "kdn_build_octree");
    const auto render_pass = symbol<decltype(&kdn_render_pass)>(lib, "kdn_render_pass");
    const auto read_estimates = symbol<decltype(&kdn_read_estimates)>(lib, "kdn_read_estimates");
    const auto write_ppm = symbol<decltype(&kdn_image_write_ppm)>(lib, "kdn_image_write_ppm");
    const auto write_pfm = symbol<decltype(&kdn_image_write_pfm)>(lib, "kdn_image_write_pfm");

    kdn_ctx* ctx = nullptr;
    auto check = [&](int rc, const char* what) {
        if (rc == KDN_OK) return;
        char message[1024] = {0};
        if (ctx) last_error(ctx, message, sizeof message);
        std::fprintf(stderr, "render_host: %s: %s\n", what, ctx ? message : global_error());
        std::exit(1);
    };

    const std::string path = argv[2], stem = argv[3];
    const int width = std::atoi(argv[4]), height = std::atoi(argv[5]);
    const uint32_t spp = (uint32_t)std::atoi(argv[6]);
    const int depth = std::atoi(argv[7]);
    std::ifstream in(path);
    std::stringstream text;
    text << in.rdbuf();
    const std::string folder = path.find('/') == std::string::npos ? "." : path.substr(0, path.rfind('/'));

    // openWorkers
    check(open(0, &ctx), "kdn_open");
    // buildWorld
    kdn_scene* scene = nullptr;
    char err[1024] = {0};
    if (scene_load(text.str().c_str(), folder.c_str(), &scene, err, sizeof err) != KDN_OK) {
        std::fprintf(stderr, "render_host: %s\n", err);
        return 1;
    }
    kdn_camera camera;
    kdn_render_settings settings;
    check(scene_settings(scene, &camera, &settings), "kdn_scene_settings");
    check(upload(ctx, scene), "kdn_upload_scene");
    float lower[3], size[3];
    check(scene_region(scene, depth, lower, size), "kdn_scene_region");
    check(build(ctx, lower, size, depth, nullptr), "kdn_build_octree");
    // drawRound: one round of spp samples, every pixel's estimate kept
    kdn_pass_desc pass{};
    pass.width = width;
    pass.height = height;
    pass.first_sample = 0;
    pass.spp = spp;
    pass.seed = (uint32_t)settings.seed;
    pass.surfaces = (uint32_t)settings.surfaces;
    pass.min_depth = 4;
    pass.flags = 1u << 6;  // the estimates
    pass.batch = settings.batch > 0 ? (uint32_t)settings.batch : spp;
    check(render_pass(ctx, &pass, nullptr), "kdn_render_pass");
    std::vector<kdn_estimate> estimates((size_t)width * height);
    check(read_estimates(ctx, estimates.data()), "kdn_read_estimates");
    // the shown frame: each pixel's mean (reconstruction/mean)
    std::vector<float> shown((size_t)width * height * 3);
    double mean = 0;
    for (size_t p = 0; p < estimates.size(); ++p) {
        const kdn_estimate& e = estimates[p];
        for (int c = 0; c < 3; ++c) {
            shown[3 * p + c] = e.taken > 0 ? (float)(e.total[c] / e.taken) : 0.0f;
            mean += shown[3 * p + c];
        }
    }
    // deliverFrame
    check(write_ppm(shown.data(), width, height, settings.exposure, settings.curve, (stem + ".ppm").c_str()), "kdn_image_write_ppm");
    check(write_pfm(shown.data(), width, height, (stem + ".pfm").c_str()), "kdn_image_write_pfm");
    // closeWorkers
    scene_free(scene);
    close(ctx);
    std::printf("render_host: %dx%d, %u spp, mean %.9g, wrote %s.ppm and %s.pfm\n", width, height, spp, mean / shown.size(),
                stem.c_str(), stem.c_str());
    dlclose(lib);
    return 0;
}
