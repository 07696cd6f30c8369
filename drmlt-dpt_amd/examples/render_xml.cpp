// Renders a scene file of the reference's XML format with integrator drmlt,
// mcmc or h2mc (lensperturbprob 0) through the C ABI of libdrmlt_hip.so, as the reference's main
// does with ParseScene and the integrator's Render (src/main.cpp:101-134):
// pre-pass, direct lighting, the chain loop with the partial outputs of
// <dpt> reportintervalseconds in <output dir>/<integrator>_partial, and the
// final film written to the film's filename (default <scene dir>/image.exr).
//
//   g++ -std=c++17 -I include render_xml.cpp -L drmlt-dpt_amd -ldrmlt_hip -o render_xml
//   ./render_xml scene.xml [iterations per chain]
#include "drmlt_amd.h"

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

static void check(int rc, const char *what) {
    if (rc < 0) {
        std::fprintf(stderr, "%s: %s\n", what, drmlt_last_error());
        std::exit(1);
    }
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s scene.xml [iterations per chain]\n", argv[0]);
        return 2;
    }
    drmlt_scene *scene = nullptr;
    check(drmlt_scene_load(argv[1], &scene), "loading the scene");
    drmlt_scene_desc desc;
    drmlt_options opt;
    check(drmlt_scene_get(scene, &desc, &opt), "reading the scene");
    std::vector<char> out(4096), integ(256);
    int32_t reportSeconds = 0;
    check(drmlt_scene_output(scene, out.data(), int32_t(out.size()), integ.data(), int32_t(integ.size()), &reportSeconds),
          "reading the output name");
    drmlt_ctx *ctx = nullptr;
    check(drmlt_ctx_create(&desc, &opt, &ctx), "creating the context");
    double b = 0;
    std::vector<float> lengthFunc(16);
    check(drmlt_prepass(ctx, &b, lengthFunc.data(), int32_t(lengthFunc.size())), "pre-pass");
    check(drmlt_direct_lighting(ctx), "direct lighting");
    const int W = int(desc.camera[10]), H = int(desc.camera[11]);
    // iterations per chain: spp * pixels / chains (delayed.cpp:206-210), or the second argument
    const int64_t iters = argc > 2 ? std::atoll(argv[2]) : int64_t(opt.spp) * W * H / opt.num_chains;
    const std::string output(out.data()), integrator(integ.data());
    const std::string dir = output.substr(0, output.find_last_of('/')) + "/" + integrator + "_partial";
    drmlt_report rep{double(reportSeconds), dir.c_str(), integrator.c_str(), DRMLT_EXR_ZIP, 1, 0};
    check(drmlt_render(ctx, iters, DRMLT_BATCHED_GRAPH, &rep), "rendering");
    std::vector<float> rgb(size_t(W) * H * 3);
    check(drmlt_read_film(ctx, rgb.data(), W, H, 0), "reading the film");
    check(drmlt_write_exr(output.c_str(), rgb.data(), W, H, DRMLT_EXR_ZIP), "writing the image");
    drmlt_stats st;
    check(drmlt_read_stats(ctx, &st), "reading the statistics");
    std::printf("b = %g, %lld iterations per chain, %llu mutations; image written to %s\n", b, (long long)iters,
                (unsigned long long)(st.large_step + st.stage1), output.c_str());
    drmlt_ctx_destroy(ctx);
    drmlt_scene_free(scene);
    return 0;
}
