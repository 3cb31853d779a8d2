// pg3gpu: command-line renderer over libpg3g, mirroring PG3Render's own command line
// (Config::ProcessCommandline, src/config.hxx:553-1098) and output (Framebuffer::SaveHDR,
// src/frame_buffer.hxx:166-200).  The GPU renderer replaces the OpenMP loop of Render()
// (src/pg3render.cxx:54-174): one context renders all iterations and the FP64 framebuffer is
// divided by the iteration count exactly as GetFramebuffer does (src/renderer.hxx:255-261).
//
//   pg3gpu [-s <scene_id>] [-w <workload>] [-a <pt|ptn|dbs|dlsa|dlss|dmis>] [-i <iterations>]
//          [-minpl <n>] [-maxpl <n>] [-iic <f>] [-sb <f>] [-em <type>] [-emfile <file>]
//          [-r <resolution>] [-seed <n>] [-k <megakernel|wavefront>] [-o <file.hdr|file.bmp>]
//          [-gamma <g>]
// -emfile names the image of an EM type 4..12 (Radiance .hdr, PFM or raw float) in place of the
// reference's OpenEXR light probe; it is rotated and scaled as the reference loads that type.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pg3g.h"
#include "pg3g_scenes.h"

namespace {

int Check(int rc, const char *what) {
    if (rc != PG3G_OK) {
        std::fprintf(stderr, "pg3gpu: %s failed (%d): %s\n", what, rc, pg3g_last_error());
        std::exit(1);
    }
    return rc;
}

}  // namespace

int main(int argc, char **argv) {
    int sceneId = -1, workload = 0, algorithm = PG3G_ALG_PATH_TRACING, kernel = PG3G_KERNEL_MEGAKERNEL;
    uint32_t iterations = 1, emType = 0, res = 0;
    long minPL = -1, maxPL = -1;
    double iic = -1, sb = -1;
    unsigned long long seed = 1234;
    std::string out = "pg3gpu.hdr";
    const char *emFile = nullptr;
    float gamma = 2.2f;
    const char *acronyms[] = {"el", "nv", "dbs", "dlsa", "dlss", "dmis", "ptn", "pt"};
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        const char *v = (i + 1 < argc) ? argv[i + 1] : nullptr;
        auto need = [&]() { if (!v) { std::fprintf(stderr, "pg3gpu: %s needs a value\n", a.c_str()); std::exit(2); } ++i; return v; };
        if (a == "-s") sceneId = std::atoi(need());
        else if (a == "-w") workload = std::atoi(need());
        else if (a == "-a") {
            const std::string s = need();
            algorithm = -1;
            for (int k = 0; k < 8; ++k) if (s == acronyms[k]) algorithm = k;
            if (algorithm < PG3G_ALG_DIRECT_BSDF) { std::fprintf(stderr, "pg3gpu: unknown or unsupported algorithm %s\n", s.c_str()); return 2; }
        } else if (a == "-i") iterations = (uint32_t)std::atol(need());
        else if (a == "-minpl") minPL = std::atol(need());
        else if (a == "-maxpl") maxPL = std::atol(need());
        else if (a == "-iic") iic = std::atof(need());
        else if (a == "-sb") sb = std::atof(need());
        else if (a == "-em") emType = (uint32_t)std::atol(need());
        else if (a == "-emfile") emFile = need();
        else if (a == "-gamma") gamma = (float)std::atof(need());
        else if (a == "-r") res = (uint32_t)std::atol(need());
        else if (a == "-seed") seed = std::strtoull(need(), nullptr, 10);
        else if (a == "-k") kernel = (std::string(need()) == "wavefront") ? PG3G_KERNEL_WAVEFRONT : PG3G_KERNEL_MEGAKERNEL;
        else if (a == "-o") out = need();
        else { std::fprintf(stderr, "pg3gpu: unknown option %s\n", a.c_str()); return 2; }
    }
    pg3g_scene_build *sb_ = nullptr;
    if (workload > 0) Check(pg3g_build_workload_scene(workload, seed, 0, res, &sb_), "pg3g_build_workload_scene");
    else if (emFile)
        Check(pg3g_build_reference_scene_em(sceneId < 0 ? 2u : (uint32_t)sceneId, emType, res, emFile, &sb_),
              "pg3g_build_reference_scene_em");
    else Check(pg3g_build_reference_scene(sceneId < 0 ? 2u : (uint32_t)sceneId, emType, res, &sb_), "pg3g_build_reference_scene");
    pg3g_scene_desc *d = pg3g_scene_build_mutable(sb_);
    pg3g_render_params &p = d->params;
    p.algorithm = algorithm;
    if (minPL >= 0) p.min_path_length = (uint32_t)minPL;
    if (maxPL >= 0) p.max_path_length = (uint32_t)maxPL;
    if (iic >= 0) p.indirect_clipping = (float)iic;
    if (sb >= 0) p.splitting_budget = (float)sb;
    p.seed = seed;

    pg3g_context *ctx = nullptr;
    Check(pg3g_create(&ctx, 0), "pg3g_create");
    Check(pg3g_set_kernel_mode(ctx, kernel), "pg3g_set_kernel_mode");
    Check(pg3g_upload_scene(ctx, d), "pg3g_upload_scene");
    const auto t0 = std::chrono::steady_clock::now();
    Check(pg3g_render(ctx, 0, iterations), "pg3g_render");
    Check(pg3g_synchronize(ctx), "pg3g_synchronize");
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const uint32_t w = d->camera.res_x, h = d->camera.res_y;
    std::vector<double> img((size_t)w * h * 3);
    uint64_t it = 0;
    Check(pg3g_read_framebuffer(ctx, img.data(), &it), "pg3g_read_framebuffer");
    for (double &v : img) v /= (double)(it ? it : 1);
    const bool ok = pg3g_save_image(out.c_str(), img.data(), w, h, 1.0, gamma) == PG3G_OK;
    if (!ok) std::fprintf(stderr, "pg3gpu: %s\n", pg3g_last_error());
    double mean = 0;
    for (double v : img) mean += v;
    mean /= (double)img.size();
    std::printf("pg3gpu: %ux%u, %u iterations, %s, %.3f s wall (%.1f Msamples/s), mean radiance %.6f -> %s\n", w, h,
                (unsigned)it, acronyms[algorithm], secs, (double)w * h * it / secs / 1e6, mean, ok ? out.c_str() : "(write failed)");
    pg3g_destroy(ctx);
    pg3g_scene_build_free(sb_);
    return ok ? 0 : 1;
}
