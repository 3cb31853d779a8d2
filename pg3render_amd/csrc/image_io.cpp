// Host-side image and environment-map files: the output step of the renderer (Framebuffer::SaveHDR
// and SaveBMP, src/frame_buffer.hxx:114-200), an EM loader without OpenEXR (the loading, azimuth
// rotation and scaling of EnvironmentMapImage::LoadImage, src/em_image.hxx:51-93, for Radiance
// RGBE, PFM and a raw float format), and the per-pixel relative-RMSE comparison that the
// reference's test script delegates to an external tool (testing/full_test.sh:97).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "image_io.hpp"

namespace pg3 {
namespace {

struct File {
    FILE *f = nullptr;
    explicit File(const char *path, const char *mode) : f(std::fopen(path, mode)) {}
    ~File() { if (f) std::fclose(f); }
};

bool EndsWith(const std::string &s, const char *suffix) {
    const size_t n = std::strlen(suffix);
    if (s.size() < n) return false;
    for (size_t i = 0; i < n; ++i)
        if (std::tolower((unsigned char)s[s.size() - n + i]) != suffix[i]) return false;
    return true;
}

// Reads one header line (without the newline); false at end of file.
bool ReadLine(FILE *f, std::string &line) {
    line.clear();
    int c;
    while ((c = std::fgetc(f)) != EOF) {
        if (c == '\n') return true;
        line.push_back((char)c);
    }
    return !line.empty();
}

float RgbeComponent(uint8_t m, uint8_t e) {          // Radiance colr_color: (m + 0.5) * 2^(e - 136)
    return e == 0 ? 0.f : (float)std::ldexp((double)m + 0.5, (int)e - 136);
}

bool ReadHdr(const char *path, std::vector<float> &rgb, uint32_t &w, uint32_t &h, std::string &err) {
    File F(path, "rb");
    if (!F.f) { err = std::string("cannot open ") + path; return false; }
    std::string line;
    if (!ReadLine(F.f, line) || line.rfind("#?", 0) != 0) { err = "not a Radiance HDR file"; return false; }
    for (;;) {                                                  // header ends at an empty line
        if (!ReadLine(F.f, line)) { err = "truncated HDR header"; return false; }
        if (line.empty()) break;
        if (line.rfind("FORMAT=", 0) == 0 && line != "FORMAT=32-bit_rle_rgbe") { err = "unsupported HDR format " + line; return false; }
    }
    if (!ReadLine(F.f, line)) { err = "missing HDR resolution"; return false; }
    unsigned hh = 0, ww = 0;
    if (std::sscanf(line.c_str(), "-Y %u +X %u", &hh, &ww) != 2 || ww == 0 || hh == 0 || ww > 65536 || hh > 65536) {
        err = "unsupported HDR orientation or size: " + line;
        return false;
    }
    w = ww;
    h = hh;
    rgb.assign((size_t)w * h * 3, 0.f);
    std::vector<uint8_t> scan((size_t)w * 4);
    for (uint32_t y = 0; y < h; ++y) {
        uint8_t head[4];
        if (std::fread(head, 1, 4, F.f) != 4) { err = "truncated HDR data"; return false; }
        if (w >= 8 && w < 32768 && head[0] == 2 && head[1] == 2 && ((uint32_t)head[2] << 8 | head[3]) == w) {
            // run-length encoded scanline: the four components one after the other
            for (int c = 0; c < 4; ++c) {
                uint32_t x = 0;
                while (x < w) {
                    const int n = std::fgetc(F.f);
                    if (n == EOF) { err = "truncated HDR run"; return false; }
                    if (n > 128) {
                        const int v = std::fgetc(F.f);
                        const uint32_t run = (uint32_t)n - 128u;
                        if (v == EOF || x + run > w) { err = "bad HDR run"; return false; }
                        for (uint32_t k = 0; k < run; ++k) scan[(size_t)(x++) * 4 + c] = (uint8_t)v;
                    } else {
                        if (n == 0 || x + (uint32_t)n > w) { err = "bad HDR run"; return false; }
                        for (int k = 0; k < n; ++k) {
                            const int v = std::fgetc(F.f);
                            if (v == EOF) { err = "truncated HDR run"; return false; }
                            scan[(size_t)(x++) * 4 + c] = (uint8_t)v;
                        }
                    }
                }
            }
        } else {                                                // flat scanline, as SaveHDR writes
            std::memcpy(scan.data(), head, 4);
            if (w > 1 && std::fread(scan.data() + 4, 1, (size_t)(w - 1) * 4, F.f) != (size_t)(w - 1) * 4) {
                err = "truncated HDR data";
                return false;
            }
        }
        for (uint32_t x = 0; x < w; ++x) {
            const uint8_t *p = &scan[(size_t)x * 4];
            float *o = &rgb[((size_t)y * w + x) * 3];
            for (int c = 0; c < 3; ++c) o[c] = RgbeComponent(p[c], p[3]);
        }
    }
    return true;
}

// PFM ("PF" RGB or "Pf" grey; negative scale = little endian; rows stored bottom to top).
bool ReadPfm(const char *path, std::vector<float> &rgb, uint32_t &w, uint32_t &h, std::string &err) {
    File F(path, "rb");
    if (!F.f) { err = std::string("cannot open ") + path; return false; }
    char kind[3] = {0, 0, 0};
    double scale = 0;
    unsigned ww = 0, hh = 0;
    if (std::fscanf(F.f, "%2s %u %u %lf", kind, &ww, &hh, &scale) != 4 || kind[0] != 'P' ||
        (kind[1] != 'F' && kind[1] != 'f') || ww == 0 || hh == 0 || ww > 65536 || hh > 65536 || scale == 0) {
        err = "not a PFM file";
        return false;
    }
    std::fgetc(F.f);                                            // the single whitespace before the data
    const int ch = kind[1] == 'F' ? 3 : 1;
    w = ww;
    h = hh;
    std::vector<float> raw((size_t)w * h * ch);
    if (std::fread(raw.data(), sizeof(float), raw.size(), F.f) != raw.size()) { err = "truncated PFM data"; return false; }
    uint16_t probe = 1;
    const bool hostLittle = *reinterpret_cast<uint8_t *>(&probe) == 1;
    if ((scale < 0) != hostLittle)
        for (float &v : raw) {
            uint8_t b[4];
            std::memcpy(b, &v, 4);
            std::swap(b[0], b[3]);
            std::swap(b[1], b[2]);
            std::memcpy(&v, b, 4);
        }
    rgb.assign((size_t)w * h * 3, 0.f);
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c)
                rgb[((size_t)y * w + x) * 3 + c] = raw[((size_t)(h - 1 - y) * w + x) * ch + (ch == 3 ? c : 0)];
    return true;
}

// Raw float image: uint32 width, uint32 height, then width*height RGB float32, top row first
// (little endian).
bool ReadRaw(const char *path, std::vector<float> &rgb, uint32_t &w, uint32_t &h, std::string &err) {
    File F(path, "rb");
    if (!F.f) { err = std::string("cannot open ") + path; return false; }
    uint32_t dims[2];
    if (std::fread(dims, 4, 2, F.f) != 2 || dims[0] == 0 || dims[1] == 0 || dims[0] > 65536 || dims[1] > 65536) {
        err = "not a raw float image";
        return false;
    }
    w = dims[0];
    h = dims[1];
    rgb.resize((size_t)w * h * 3);
    if (std::fread(rgb.data(), sizeof(float), rgb.size(), F.f) != rgb.size()) { err = "truncated raw image"; return false; }
    return true;
}

}  // namespace

bool ReadImageFile(const char *path, std::vector<float> &rgb, uint32_t &w, uint32_t &h, std::string &err) {
    File F(path, "rb");
    if (!F.f) { err = std::string("cannot open ") + path; return false; }
    char magic[2] = {0, 0};
    const size_t n = std::fread(magic, 1, 2, F.f);
    std::fclose(F.f);
    F.f = nullptr;
    if (n == 2 && magic[0] == '#' && magic[1] == '?') return ReadHdr(path, rgb, w, h, err);
    if (n == 2 && magic[0] == 'P' && (magic[1] == 'F' || magic[1] == 'f')) return ReadPfm(path, rgb, w, h, err);
    return ReadRaw(path, rgb, w, h, err);
}

bool LoadEmFile(const char *path, float azimuthRotation, float scale, std::vector<float> &rgb, uint32_t &w,
                uint32_t &h, std::string &err) {
    std::vector<float> src;
    if (!ReadImageFile(path, src, w, h, err)) return false;
    float rot = std::fmod(azimuthRotation, 1.0f);               // Math::FmodX, math.hxx:130-139
    if (rot < 0.f) rot += 1.0f;
    const int32_t iRot = (int32_t)(rot * (float)w);
    rgb.assign(src.size(), 0.f);
    for (uint32_t j = 0; j < h; ++j)
        for (uint32_t i = 0; i < w; ++i) {
            int32_t x = (int32_t)i + iRot;
            if (x >= (int32_t)w) x -= (int32_t)w;
            for (int c = 0; c < 3; ++c)
                rgb[((size_t)j * w + (uint32_t)x) * 3 + c] = src[((size_t)j * w + i) * 3 + c] * scale;
        }
    return true;
}

bool WriteHdr(const char *path, const double *rgb, uint32_t w, uint32_t h, double scale, std::string &err) {
    File F(path, "wb");
    if (!F.f) { err = std::string("cannot write ") + path; return false; }
    std::fprintf(F.f, "#?RADIANCE\n# PG3Render\nFORMAT=32-bit_rle_rgbe\n\n-Y %u +X %u\n", h, w);
    std::vector<uint8_t> row((size_t)w * 4);
    for (uint32_t y = 0; y < h; ++y) {
        for (uint32_t x = 0; x < w; ++x) {
            const double *p = rgb + ((size_t)y * w + x) * 3;
            const double r = p[0] * scale, g = p[1] * scale, b = p[2] * scale;
            const double v = std::max(r, std::max(g, b));
            uint8_t *o = &row[(size_t)x * 4];
            o[0] = o[1] = o[2] = o[3] = 0;
            if (v >= 1e-32) {
                int e;
                const double m = std::frexp(v, &e) * 256.0 / v;
                o[0] = (uint8_t)(r * m);
                o[1] = (uint8_t)(g * m);
                o[2] = (uint8_t)(b * m);
                o[3] = (uint8_t)(e + 128);
            }
        }
        if (std::fwrite(row.data(), 1, row.size(), F.f) != row.size()) { err = "write failed"; return false; }
    }
    return true;
}

bool WriteBmp(const char *path, const double *rgb, uint32_t w, uint32_t h, double scale, float gamma, std::string &err) {
    File F(path, "wb");
    if (!F.f) { err = std::string("cannot write ") + path; return false; }
    const uint32_t rowBytes = (w * 3u + 3u) & ~3u;              // rows padded to 4 bytes
    uint8_t hd[54] = {'B', 'M'};
    auto put32 = [&](int off, uint32_t v) { for (int k = 0; k < 4; ++k) hd[off + k] = (uint8_t)(v >> (8 * k)); };
    auto put16 = [&](int off, uint16_t v) { hd[off] = (uint8_t)v; hd[off + 1] = (uint8_t)(v >> 8); };
    put32(2, 54u + rowBytes * h);
    put32(10, 54u);
    put32(14, 40u);
    put32(18, w);
    put32(22, h);
    put16(26, 1);
    put16(28, 24);
    put32(34, rowBytes * h);
    put32(38, 2953u);                                           // 75 dpi, as the reference
    put32(42, 2953u);
    if (std::fwrite(hd, 1, 54, F.f) != 54) { err = "write failed"; return false; }
    const double invGamma = 1.0 / (double)(gamma > 0.f ? gamma : 1.f);
    std::vector<uint8_t> row(rowBytes, 0);
    for (uint32_t y = 0; y < h; ++y) {
        const double *src = rgb + (size_t)(h - 1 - y) * w * 3;  // bottom row first
        for (uint32_t x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c) {
                const double v = std::pow(std::max(0.0, src[(size_t)x * 3 + (2 - c)] * scale), invGamma) * 255.0;
                row[(size_t)x * 3 + c] = (uint8_t)std::min(255.0, std::max(0.0, v));
            }
        if (std::fwrite(row.data(), 1, rowBytes, F.f) != rowBytes) { err = "write failed"; return false; }
    }
    return true;
}

bool WriteImageFile(const char *path, const double *rgb, uint32_t w, uint32_t h, double scale, float gamma,
                    std::string &err) {
    const std::string p(path);
    if (EndsWith(p, ".bmp")) return WriteBmp(path, rgb, w, h, scale, gamma, err);
    if (EndsWith(p, ".hdr")) return WriteHdr(path, rgb, w, h, scale, err);
    err = "unknown image extension (expected .hdr or .bmp): " + p;
    return false;
}

}  // namespace pg3
