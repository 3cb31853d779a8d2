// pg3diff: compares a rendered image with a reference image, the check of PG3Render's regression
// script (testing/full_test.sh:80-110), which runs an external diff tool in relative-RMSE mode on
// the two HDR outputs and counts the test as passed when the error is under a threshold.
//
//   pg3diff [-threshold <t>] [-output <diff.hdr|diff.bmp>] <rendered> <reference>
//
// Images: Radiance .hdr, PFM or raw float (pg3g_load_image).  Prints the per-pixel and whole-image
// relative RMSE of luminance (pg3g_image_compare); exit status 0 if the per-pixel relative RMSE
// is at most the threshold (default 0.05), 1 if not, 2 on errors.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "pg3g.h"

namespace {

bool Load(const char *path, std::vector<float> &rgb, uint32_t &w, uint32_t &h) {
    if (pg3g_load_image(path, &w, &h, nullptr, 0) != PG3G_OK) return false;
    rgb.resize((size_t)w * h * 3);
    return pg3g_load_image(path, &w, &h, rgb.data(), rgb.size()) == PG3G_OK;
}

}  // namespace

int main(int argc, char **argv) {
    double threshold = 0.05;
    const char *output = nullptr;
    std::vector<const char *> files;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if ((a == "-threshold" || a == "-output") && i + 1 < argc) {
            if (a == "-threshold") threshold = std::atof(argv[++i]);
            else output = argv[++i];
        } else if (!a.empty() && a[0] == '-') {
            std::fprintf(stderr, "pg3diff: unknown option %s\n", a.c_str());
            return 2;
        } else {
            files.push_back(argv[i]);
        }
    }
    if (files.size() != 2) {
        std::fprintf(stderr, "usage: pg3diff [-threshold t] [-output diff.hdr] rendered reference\n");
        return 2;
    }
    std::vector<float> a, b;
    uint32_t wa = 0, ha = 0, wb = 0, hb = 0;
    if (!Load(files[0], a, wa, ha) || !Load(files[1], b, wb, hb)) {
        std::fprintf(stderr, "pg3diff: %s\n", pg3g_last_error());
        return 2;
    }
    if (wa != wb || ha != hb) {
        std::fprintf(stderr, "pg3diff: sizes differ (%ux%u vs %ux%u)\n", wa, ha, wb, hb);
        return 2;
    }
    std::vector<float> diff(a.size());
    pg3g_image_diff d;
    if (pg3g_image_compare(a.data(), b.data(), wa, ha, &d, diff.data()) != PG3G_OK) {
        std::fprintf(stderr, "pg3diff: %s\n", pg3g_last_error());
        return 2;
    }
    if (output) {
        std::vector<double> dd(diff.begin(), diff.end());
        if (pg3g_save_image(output, dd.data(), wa, ha, 1.0, 1.0f) != PG3G_OK) {
            std::fprintf(stderr, "pg3diff: %s\n", pg3g_last_error());
            return 2;
        }
    }
    const bool pass = d.pixel_rel_rmse <= threshold;
    std::printf("pg3diff: %ux%u  pixel relative RMSE %.6f  image relative RMSE %.6f  max pixel %.4f  "
                "mean luminance %.6g / %.6g  -> %s\n",
                wa, ha, d.pixel_rel_rmse, d.image_rel_rmse, d.max_pixel_rel, d.mean_luminance,
                d.mean_luminance_reference, pass ? "passed" : "FAILED");
    return pass ? 0 : 1;
}
