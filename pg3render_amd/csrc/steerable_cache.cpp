// On-disk cache of steerable sampler tables, keyed by the content of the EM image.  The reference
// caches its built tree next to the EM file, keyed by the file name and the build parameters
// (EnvironmentMapSteerableSampler::GenerateSaveFilePath / SaveToDisk / LoadFromDisk,
// src/em_steerable_sampler.hxx:1596-2030), so an edited image under the same name reuses a stale
// tree and an image built in memory is never cached.  Here the key is a 64-bit hash of the texels
// plus the image size, the filter and every build parameter; a file that does not match in every
// field, or whose payload checksum fails, is ignored and rebuilt.  Files are written to a
// temporary name and renamed, so concurrent writers never expose a partial file.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unistd.h>
#include <vector>

#include "pg3g_internal.hpp"

namespace pg3 {
namespace {

constexpr char kMagic[] = "PG3G steerable tables v1\n";
constexpr uint32_t kSamplesPerDim = 50, kChildren = 6;      // build constants recorded in the file
constexpr float kOversampling = 0.7f, kMaxSpan = 1.1f;

std::mutex gDirMutex;
std::string gDir;
bool gDirSet = false;

// FNV-1a over 64-bit words (then the tail bytes): a content hash, not a cryptographic one.
uint64_t Hash64(const void *data, size_t bytes, uint64_t h = 0xcbf29ce484222325ull) {
    const uint8_t *p = static_cast<const uint8_t *>(data);
    size_t i = 0;
    for (; i + 8 <= bytes; i += 8) {
        uint64_t v;
        std::memcpy(&v, p + i, 8);
        h = (h ^ v) * 0x100000001b3ull;
    }
    for (; i < bytes; ++i) h = (h ^ p[i]) * 0x100000001b3ull;
    return h;
}

struct Key {
    uint64_t hash;
    uint32_t w, h, bilinear, minLevel, maxLevel, samplesPerDim, children;
    float maxErr, oversampling, maxSpan;
};

std::string CacheDir() {
    std::lock_guard<std::mutex> lock(gDirMutex);
    if (gDirSet) return gDir;
    const char *e = std::getenv("PG3G_STEERABLE_CACHE_DIR");
    return e ? std::string(e) : std::string();
}

std::string FileName(const std::string &dir, const Key &k) {
    char buf[160];
    std::snprintf(buf, sizeof(buf), "/%016llx_%ux%u_%s_e%.2f_sll%u_slu%u.pg3st", (unsigned long long)k.hash, k.w, k.h,
                  k.bilinear ? "bi" : "nn", (double)k.maxErr, k.minLevel, k.maxLevel);
    return dir + buf;
}

template <class T> bool Put(FILE *f, const T *v, size_t n) { return std::fwrite(v, sizeof(T), n, f) == n; }
template <class T> bool Get(FILE *f, T *v, size_t n) { return std::fread(v, sizeof(T), n, f) == n; }

uint64_t PayloadHash(const SteerableTables &t) {
    uint64_t h = Hash64(t.levelCount.data(), t.levelCount.size() * 4);
    h = Hash64(t.levelBase.data(), t.levelBase.size() * 4, h);
    h = Hash64(t.nodeW.data(), t.nodeW.size() * 4, h);
    h = Hash64(t.leafVerts.data(), t.leafVerts.size() * 4, h);
    return Hash64(t.rootW, sizeof(t.rootW), h);
}

bool Load(const std::string &path, const Key &key, SteerableTables &t) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    bool ok = false;
    do {
        char magic[sizeof(kMagic)] = {};
        Key k;
        uint32_t counts[3];
        uint64_t sizes[2], check;
        if (!Get(f, magic, sizeof(kMagic) - 1) || std::memcmp(magic, kMagic, sizeof(kMagic) - 1) != 0) break;
        if (!Get(f, &k, 1) || std::memcmp(&k, &key, sizeof(Key)) != 0) break;
        if (!Get(f, counts, 3)) break;
        t.numLeaves = counts[0];
        t.numVertices = counts[1];
        t.numLevels = counts[2];
        if (t.numLevels == 0 || t.numLevels > (uint32_t)kMaxTreeLevels) break;
        t.levelCount.resize(t.numLevels);
        t.levelBase.resize(t.numLevels);
        if (!Get(f, t.levelCount.data(), t.numLevels) || !Get(f, t.levelBase.data(), t.numLevels)) break;
        if (!Get(f, t.rootW, 9) || !Get(f, sizes, 2)) break;
        uint64_t nodes = 0;
        bool shape = t.levelCount[0] == t.numLeaves && t.levelCount[t.numLevels - 1] == 1;
        for (uint32_t l = 0; l < t.numLevels && shape; ++l) {
            shape = t.levelBase[l] == nodes;
            nodes += t.levelCount[l];
        }
        if (!shape || sizes[0] != nodes * 12u || sizes[1] != (uint64_t)t.numLeaves * 36u) break;
        t.nodeW.resize(sizes[0]);
        t.leafVerts.resize(sizes[1]);
        if (!Get(f, t.nodeW.data(), sizes[0]) || !Get(f, t.leafVerts.data(), sizes[1]) || !Get(f, &check, 1)) break;
        ok = check == PayloadHash(t);
    } while (false);
    std::fclose(f);
    return ok;
}

void Save(const std::string &path, const Key &key, const SteerableTables &t) {
    const std::string tmp = path + ".tmp." + std::to_string((long)getpid());
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) return;
    const uint32_t counts[3] = {t.numLeaves, t.numVertices, t.numLevels};
    const uint64_t sizes[2] = {t.nodeW.size(), t.leafVerts.size()};
    const uint64_t check = PayloadHash(t);
    const bool ok = Put(f, kMagic, sizeof(kMagic) - 1) && Put(f, &key, 1) && Put(f, counts, 3) &&
                    Put(f, t.levelCount.data(), t.numLevels) && Put(f, t.levelBase.data(), t.numLevels) &&
                    Put(f, t.rootW, 9) && Put(f, sizes, 2) && Put(f, t.nodeW.data(), t.nodeW.size()) &&
                    Put(f, t.leafVerts.data(), t.leafVerts.size()) && Put(f, &check, 1);
    const bool closed = std::fclose(f) == 0;
    if (ok && closed) std::rename(tmp.c_str(), path.c_str());
    else std::remove(tmp.c_str());
}

}  // namespace

void SetSteerableCacheDir(const char *dir) {
    std::lock_guard<std::mutex> lock(gDirMutex);
    gDirSet = true;
    gDir = dir ? std::string(dir) : std::string();
}

bool BuildSteerableCached(const float *texels, uint32_t w, uint32_t h, bool bilinear, float maxErr, uint32_t minLevel,
                          uint32_t maxLevel, SteerableTables &out, std::string &err, uint32_t threads, bool *fromCache) {
    if (fromCache) *fromCache = false;
    const std::string dir = CacheDir();
    if (dir.empty() || !texels || w == 0 || h == 0)
        return BuildSteerable(texels, w, h, bilinear, maxErr, minLevel, maxLevel, out, err, threads);
    Key key{};
    key.hash = Hash64(texels, (size_t)w * h * 3 * sizeof(float));
    key.w = w;
    key.h = h;
    key.bilinear = bilinear ? 1u : 0u;
    key.minLevel = minLevel;
    key.maxLevel = maxLevel;
    key.samplesPerDim = kSamplesPerDim;
    key.children = kChildren;
    key.maxErr = maxErr > 0.f ? maxErr : 5.f;
    key.oversampling = kOversampling;
    key.maxSpan = kMaxSpan;
    const std::string path = FileName(dir, key);
    if (Load(path, key, out)) {
        if (fromCache) *fromCache = true;
        return true;
    }
    out = SteerableTables{};
    if (!BuildSteerable(texels, w, h, bilinear, maxErr, minLevel, maxLevel, out, err, threads)) return false;
    Save(path, key, out);
    return true;
}

}  // namespace pg3
