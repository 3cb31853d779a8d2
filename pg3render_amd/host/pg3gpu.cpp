// pg3gpu: command-line renderer over libpg3g, mirroring PG3Render's own command line
// (Config::ProcessCommandline, src/config.hxx:553-1098), its default output file name
// (Config::DefaultFilename, :172-264) and output (Framebuffer::SaveHDR / SaveBMP,
// src/frame_buffer.hxx:114-200).  The GPU renderer replaces the OpenMP loop of Render()
// (src/pg3render.cxx:54-174): one context renders all iterations (or, with -t, as many as fit in
// the time) and the FP64 framebuffer is divided by the iteration count exactly as GetFramebuffer
// does (src/renderer.hxx:255-261).
//
// The reference's options: -s <scene_id> -a <algorithm> -t <time> | -i <iterations>
//   -minpl <n> -maxpl <n> -iic <f> -sb|--splitting-budget <n> -sl <level>
//   -slbr|--splitting-light-to-bsdf-ratio <f> -em <env_map_type> -e <bmp|hdr> -od <dir> -o <name>
//   -ot <trail> -j <threads> -q -opop|--only-print-output-pathname
//   -opof|--only-print-output-filename -h|--help|/?
// GPU-only options: -w <workload> (benchmark scenes), -r <resolution>, -seed <n>,
//   -k <megakernel|wavefront>, -emfile <file> (the image of an EM type 4..12: Radiance .hdr, PFM
//   or raw float, rotated and scaled as the reference loads that type, in place of its OpenEXR
//   light probe), -gamma <g> (BMP output).
// -j is accepted and checked; the GPU renders with all its compute units.  The output directory
// is joined with '/' (the reference joins with '\').
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
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

void PrintHelp(const char *argv0) {
    std::printf(
        "\nUsage: %s [-s <scene_id>] [-a <algorithm>] [-t <time> | -i <iterations>] "
        "[-minpl <min_path_length>] [-maxpl <max_path_length>] [-iic <indirect_illum_clipping_value>] "
        "[-sb|--splitting-budget <splitting_budget>] [-sl <splitting_level>] "
        "[-slbr|--splitting-light-to-bsdf-ratio <splitting_light_to_bsdf_ratio>] [-em <env_map_type>] "
        "[-e <def_output_ext>] [-od <output_directory>] [-o <output_name>] [-ot <output_trail>] "
        "[-j <threads_count>] [-q] [-opop|--only-print-output-pathname] [-opof|--only-print-output-filename] "
        "[-w <workload>] [-r <resolution>] [-seed <n>] [-k <megakernel|wavefront>] [-emfile <file>] "
        "[-gamma <g>]\n\n"
        "    -a     algorithm: el, dbs, dlsa, dlss, dmis, ptn, pt (el and nv are not on the GPU)\n"
        "    -t     render for <time> seconds (takes precedence over -i)\n"
        "    -maxpl 0 = Russian roulette\n\n",
        argv0);
}

// Fixed-point text as the reference's std::ostringstream with precision(1) and std::fixed.
std::string Fixed1(double v) {
    std::ostringstream s;
    s.precision(1);
    s << std::fixed << v;
    return s.str();
}

}  // namespace

int main(int argc, char **argv) {
    const char *acronyms[] = {"el", "nv", "dbs", "dlsa", "dlss", "dmis", "ptn", "pt"};
    int sceneId = 0, workload = 0, algorithm = PG3G_ALG_PATH_TRACING, kernel = PG3G_KERNEL_MEGAKERNEL;
    uint32_t emType = 0, res = 0, threads = 0;
    long iterations = 1, minPL = 1, maxPL = 0;
    double maxTime = -1, iic = 16, splitLevel = 1, lightRatio = 1;
    long splitBudget = 4;
    unsigned long long seed = 1234;
    std::string outName, outDir, trail, ext = "bmp";
    const char *emFile = nullptr;
    float gamma = 2.2f;
    bool quiet = false;
    int onlyPrint = 0;   // 1: output path name, 2: output file name
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-h" || a == "--help" || a == "/?") {
            PrintHelp(argv[0]);
            return 1;
        }
    }
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a.empty() || a[0] != '-') continue;   // the reference skips words that are not options
        const char *v = (i + 1 < argc) ? argv[i + 1] : nullptr;
        auto need = [&]() {
            if (!v) { std::printf("Error: Missing argument of %s, please see help (-h)\n", a.c_str()); std::exit(1); }
            ++i;
            return v;
        };
        auto bad = [&](const char *what) {
            std::printf("Error: Invalid <%s> argument \"%s\", please see help (-h)\n", what, argv[i]);
            std::exit(1);
        };
        // numbers as the reference reads them (std::istringstream >>): a value that does not
        // start with a number is an error
        auto num = [&](const char *what) {
            const char *t = need();
            char *end = nullptr;
            const double x = std::strtod(t, &end);
            if (end == t) bad(what);
            return x;
        };
        if (a == "-opop" || a == "--only-print-output-pathname") onlyPrint = 1;
        else if (a == "-opof" || a == "--only-print-output-filename") onlyPrint = 2;
        else if (a == "-j") { const long n = (long)num("threads_count"); if (n <= 0) bad("threads_count"); threads = (uint32_t)n; }
        else if (a == "-q") quiet = true;
        else if (a == "-s") { sceneId = (int)num("scene_id"); if (sceneId < 0) bad("scene_id"); }
        else if (a == "-w") workload = (int)num("workload");
        else if (a == "-a") {
            const std::string s = need();
            algorithm = -1;
            for (int k = 0; k < 8; ++k) if (s == acronyms[k]) algorithm = k;
            if (algorithm < 0) bad("algorithm");
            if (algorithm < PG3G_ALG_DIRECT_BSDF) {
                std::fprintf(stderr, "pg3gpu: algorithm %s is not on the GPU path\n", s.c_str());
                return 2;
            }
        } else if (a == "-t") { maxTime = num("time"); if (maxTime < 0) bad("time"); iterations = -1; }
        else if (a == "-i") { iterations = (long)num("iterations"); if (iterations < 1) bad("iterations"); maxTime = -1; }
        else if (a == "-minpl") { minPL = (long)num("min_path_length"); if (minPL < 0) bad("min_path_length"); }
        else if (a == "-maxpl") { maxPL = (long)num("max_path_length"); if (maxPL < 0) bad("max_path_length"); }
        else if (a == "-iic") { iic = num("indirect_illum_clipping_value"); if (iic < 0) bad("indirect_illum_clipping_value"); }
        else if (a == "-sb" || a == "--splitting-budget") { splitBudget = (long)num("splitting_budget"); if (splitBudget < 1) bad("splitting_budget"); }
        else if (a == "-sl") { splitLevel = num("splitting_level"); if (!(splitLevel > 0 && splitLevel <= 1)) bad("splitting_level"); }
        else if (a == "-slbr" || a == "--splitting-light-to-bsdf-ratio") {
            lightRatio = num("splitting_light_to_bsdf_ratio");
            if (!(lightRatio > 0)) bad("splitting_light_to_bsdf_ratio");
        } else if (a == "-em") { const double e = num("env_map_type"); if (e < 0 || e > 12) bad("env_map_type"); emType = (uint32_t)e; }
        else if (a == "-e") {
            ext = need();
            if (ext != "bmp" && ext != "hdr") {
                std::printf("\nWarning: The <default_output_extension> argument \"%s\" is neither \"bmp\" nor \"hdr\". "
                            "Using \"bmp\".\n\n", ext.c_str());
                ext = "bmp";
            }
        } else if (a == "-o") { outName = need(); if (outName.empty()) bad("output_name"); }
        else if (a == "-od") { outDir = need(); if (outDir.empty()) bad("output_directory"); }
        else if (a == "-ot") { const std::string t = need(); if (t.empty()) bad("output_trail"); trail += t; }
        else if (a == "-emfile") emFile = need();
        else if (a == "-gamma") gamma = (float)num("gamma");
        else if (a == "-r") res = (uint32_t)num("resolution");
        else if (a == "-seed") seed = std::strtoull(need(), nullptr, 10);
        else if (a == "-k") kernel = (std::string(need()) == "wavefront") ? PG3G_KERNEL_WAVEFRONT : PG3G_KERNEL_MEGAKERNEL;
        else { std::printf("Error: Unknown option \"%s\", please see help (-h)\n", a.c_str()); return 1; }
    }
    (void)threads;

    // Config::DefaultFilename (config.hxx:172-264).  Reference scenes use the reference's default
    // EM sampler (simple spherical: no suffix), so their name needs no scene build.
    auto defaultName = [&](const char *acr, int alg, uint32_t pmin, uint32_t pmax, double budget, double level,
                           double ratio, double clip, const char *samplerSuffix) {
        std::string n = std::string(acr) + "_" + acronyms[alg];
        if (alg == PG3G_ALG_PATH_TRACING_NAIVE || alg == PG3G_ALG_PATH_TRACING)
            n += (pmax == 0) ? std::string("_rr") : "_pl" + std::to_string(pmin) + "-" + std::to_string(pmax);
        if (alg == PG3G_ALG_PATH_TRACING) {
            n += "_splt" + std::to_string((long)budget) + "," + Fixed1(level) + "," + Fixed1(ratio);
            if (clip > 0.0) n += "_iic" + Fixed1(clip);
        }
        n += samplerSuffix;
        n += "_";
        if (maxTime > 0) {
            std::ostringstream t;
            t.width(2);
            t.fill('0');
            t << maxTime;
            n += t.str() + "sec";
        } else {
            n += std::to_string(iterations) + "s";
        }
        if (!trail.empty()) n += "_" + trail;
        return n + "." + ext;
    };
    auto finishName = [&]() {
        const std::string tail = outName.size() > 4 ? outName.substr(outName.size() - 4) : std::string();
        if (tail != ".bmp" && tail != ".hdr") outName += ".bmp";
        return outDir.empty() ? outName : outDir + "/" + outName;
    };
    if (workload == 0 && outName.empty()) {
        char acr[128];
        Check(pg3g_reference_scene_acronym((uint32_t)sceneId, emType, acr, sizeof(acr)), "pg3g_reference_scene_acronym");
        outName = defaultName(acr, algorithm, (uint32_t)minPL, (uint32_t)maxPL, (double)splitBudget, splitLevel,
                              lightRatio, iic, "");
    }
    if (onlyPrint && workload == 0) {
        const std::string full = finishName();
        std::printf("%s", onlyPrint == 1 ? full.c_str() : outName.c_str());
        return 1;
    }

    pg3g_scene_build *sb_ = nullptr;
    if (workload > 0) Check(pg3g_build_workload_scene(workload, seed, 0, res, &sb_), "pg3g_build_workload_scene");
    else if (emFile) Check(pg3g_build_reference_scene_em((uint32_t)sceneId, emType, res, emFile, &sb_), "pg3g_build_reference_scene_em");
    else Check(pg3g_build_reference_scene((uint32_t)sceneId, emType, res, &sb_), "pg3g_build_reference_scene");
    pg3g_scene_desc *d = pg3g_scene_build_mutable(sb_);
    pg3g_render_params &p = d->params;
    if (workload == 0) {
        // the reference's settings (workload scenes keep their own)
        p.algorithm = algorithm;
        p.min_path_length = (uint32_t)minPL;
        p.max_path_length = (uint32_t)maxPL;
        p.indirect_clipping = (float)iic;
        p.splitting_budget = (float)splitBudget;
        p.splitting_level = (float)splitLevel;
        p.light_to_bsdf_ratio = (float)lightRatio;
    } else {
        p.algorithm = algorithm;
    }
    p.seed = seed;

    if (outName.empty()) {   // benchmark workloads: "w<n>" and the workload's own settings and sampler
        char acr[16];
        std::snprintf(acr, sizeof(acr), "w%d", workload);
        const char *suffix = (d->em.width > 0 && d->em.sampler == PG3G_EM_STEERABLE) ? "_emss"
                             : (d->em.width > 0 && d->em.sampler == PG3G_EM_COSINE) ? "_emcs" : "";
        outName = defaultName(acr, p.algorithm, p.min_path_length, p.max_path_length, p.splitting_budget,
                              p.splitting_level, p.light_to_bsdf_ratio, p.indirect_clipping, suffix);
    }
    const std::string fullPath = finishName();
    if (onlyPrint) {
        std::printf("%s", onlyPrint == 1 ? fullPath.c_str() : outName.c_str());
        pg3g_scene_build_free(sb_);
        return 1;
    }

    pg3g_context *ctx = nullptr;
    Check(pg3g_create(&ctx, 0), "pg3g_create");
    Check(pg3g_set_kernel_mode(ctx, kernel), "pg3g_set_kernel_mode");
    Check(pg3g_upload_scene(ctx, d), "pg3g_upload_scene");
    const auto t0 = std::chrono::steady_clock::now();
    auto elapsed = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
    uint64_t done = 0;
    if (maxTime > 0) {
        // time-based loop (pg3render.cxx:73-100): render chunks sized to a tenth of the time left
        // at the measured rate until the time is up
        uint32_t chunk = 1;
        while (elapsed() < maxTime) {
            Check(pg3g_render(ctx, (uint32_t)done, chunk), "pg3g_render");
            Check(pg3g_synchronize(ctx), "pg3g_synchronize");
            done += chunk;
            const double rate = done / std::max(elapsed(), 1e-6), left = maxTime - elapsed();
            chunk = (uint32_t)std::max(1.0, std::min(1e6, std::floor(rate * left * 0.1)));
        }
    } else {
        Check(pg3g_render(ctx, 0, (uint32_t)iterations), "pg3g_render");
        Check(pg3g_synchronize(ctx), "pg3g_synchronize");
    }
    const double secs = elapsed();
    const uint32_t w = d->camera.res_x, h = d->camera.res_y;
    std::vector<double> img((size_t)w * h * 3);
    uint64_t it = 0;
    Check(pg3g_read_framebuffer(ctx, img.data(), &it), "pg3g_read_framebuffer");
    for (double &v : img) v /= (double)(it ? it : 1);
    const bool ok = pg3g_save_image(fullPath.c_str(), img.data(), w, h, 1.0, gamma) == PG3G_OK;
    if (!ok) std::fprintf(stderr, "pg3gpu: %s\n", pg3g_last_error());
    if (!quiet) {
        double mean = 0;
        for (double v : img) mean += v;
        mean /= (double)img.size();
        std::printf("pg3gpu: %ux%u, %llu iterations, %s, %.3f s wall (%.1f Msamples/s), mean radiance %.6f -> %s\n", w,
                    h, (unsigned long long)it, acronyms[p.algorithm], secs, (double)w * h * it / secs / 1e6, mean,
                    ok ? fullPath.c_str() : "(write failed)");
    }
    pg3g_destroy(ctx);
    pg3g_scene_build_free(sb_);
    return ok ? 0 : 1;
}
