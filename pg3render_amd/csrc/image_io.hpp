// Image files (image_io.cpp): Radiance RGBE, PFM or raw float reads; .hdr / .bmp writes of an RGB
// FP64 image times `scale`.  Host only.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace pg3 {

bool ReadImageFile(const char *path, std::vector<float> &rgb, uint32_t &w, uint32_t &h, std::string &err);
// EnvironmentMapImage::LoadImage (src/em_image.hxx:51-93) on ReadImageFile: azimuth rotation by a
// fraction of a turn (modulo 1) and a scale.
bool LoadEmFile(const char *path, float azimuthRotation, float scale, std::vector<float> &rgb, uint32_t &w,
                uint32_t &h, std::string &err);
bool WriteImageFile(const char *path, const double *rgb, uint32_t w, uint32_t h, double scale, float gamma,
                    std::string &err);

}  // namespace pg3
