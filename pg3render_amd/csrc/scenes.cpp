// Host scene builders: the benchmark workloads (seeded Cornell-box variants and synthetic
// environment maps) and the reference's own Cornell-box scene IDs (src/scene.hxx:176-956,
// src/config.hxx:45-115), all emitted as pg3g_scene_desc.
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "pg3g.h"
#include "pg3g_scenes.h"
#include "image_io.hpp"

struct pg3g_scene_build {
    std::vector<float> tris, sphs, em;
    std::vector<int32_t> triMat, sphMat;
    std::vector<pg3g_material> mats;
    std::vector<pg3g_light> lights;
    pg3g_scene_desc desc{};
    uint32_t spp = 1;

    void Finish() {
        desc.num_triangles = (uint32_t)triMat.size();
        desc.triangles = tris.data();
        desc.triangle_mat = triMat.data();
        desc.num_spheres = (uint32_t)sphMat.size();
        desc.spheres = sphs.data();
        desc.sphere_mat = sphMat.data();
        desc.num_materials = (uint32_t)mats.size();
        desc.materials = mats.data();
        desc.num_lights = (uint32_t)lights.size();
        desc.lights = lights.data();
        desc.em.texels = em.empty() ? nullptr : em.data();
    }
};

namespace {

const float kPi = 3.14159265358979f;

struct Rand {       // splitmix64 stream for seeded scene parameters
    uint64_t s;
    uint64_t Next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double U() { return (double)(Next() >> 11) * (1.0 / 9007199254740992.0); }
    double U(double a, double b) { return a + (b - a) * U(); }
};

pg3g_material Mat(int type) {
    pg3g_material m;
    std::memset(&m, 0, sizeof(m));
    m.type = type;
    m.light_id = -1;
    m.glossy[0] = m.glossy[1] = m.glossy[2] = 1.f;
    m.phong_exponent = 1.f;
    m.alpha = 0.1f;
    m.inner_ior = 1.f;
    m.outer_ior = 1.f;
    return m;
}
pg3g_material Lambert(float r, float g, float b) {
    pg3g_material m = Mat(PG3G_MAT_LAMBERT);
    m.diffuse[0] = r; m.diffuse[1] = g; m.diffuse[2] = b;
    m.glossy[0] = m.glossy[1] = m.glossy[2] = 0.f;
    return m;
}
// Complex IOR (eta, k) with normal-incidence reflectance F0 for a chosen eta:
// F0 = ((eta-1)^2 + k^2) / ((eta+1)^2 + k^2)  (physics.hxx:110-141 at cos = 1)
float ConductorK(float eta, float f0) {
    const double k2 = (f0 * (eta + 1.0) * (eta + 1.0) - (eta - 1.0) * (eta - 1.0)) / (1.0 - f0);
    return (float)std::sqrt(std::max(k2, 0.0));
}
pg3g_material GgxConductor(float alpha, float f0) {
    pg3g_material m = Mat(PG3G_MAT_GGX_CONDUCTOR);
    m.alpha = alpha;
    m.inner_ior = 0.2f;
    m.outer_ior = 1.f;
    m.absorbance = ConductorK(0.2f, f0);
    return m;
}
pg3g_material GgxDielectric(float alpha, float ior) {
    pg3g_material m = Mat(PG3G_MAT_GGX_DIELECTRIC);
    m.alpha = alpha;
    m.inner_ior = ior;
    m.outer_ior = 1.f;
    return m;
}

void AddTri(pg3g_scene_build &b, const float p0[3], const float p1[3], const float p2[3], int mat) {
    b.tris.insert(b.tris.end(), p0, p0 + 3);
    b.tris.insert(b.tris.end(), p1, p1 + 3);
    b.tris.insert(b.tris.end(), p2, p2 + 3);
    b.triMat.push_back(mat);
}
void AddQuad(pg3g_scene_build &b, const float a[3], const float c[3], const float d[3], const float e[3], int mat) {
    AddTri(b, a, c, d, mat);   // (a, c, d) and (a, d, e): both wound like the quad
    AddTri(b, a, d, e, mat);
}
void AddSphere(pg3g_scene_build &b, float x, float y, float z, float r, int mat) {
    const float s[4] = {x, y, z, r};
    b.sphs.insert(b.sphs.end(), s, s + 4);
    b.sphMat.push_back(mat);
}
pg3g_light AreaLight(const float p0[3], const float p1[3], const float p2[3], float power) {
    pg3g_light l;
    std::memset(&l, 0, sizeof(l));
    l.type = PG3G_LIGHT_AREA;
    for (int k = 0; k < 3; ++k) { l.p0[k] = p0[k]; l.p1[k] = p1[k]; l.p2[k] = p2[k]; l.power[k] = power; }
    return l;
}
pg3g_light Background(float r, float g, float b) {
    pg3g_light l;
    std::memset(&l, 0, sizeof(l));
    l.type = PG3G_LIGHT_BACKGROUND;
    l.power[0] = r; l.power[1] = g; l.power[2] = b;
    return l;
}
void SetParams(pg3g_scene_build &b, uint32_t maxPL, uint64_t seed, uint32_t sceneId) {
    pg3g_render_params &p = b.desc.params;
    std::memset(&p, 0, sizeof(p));
    p.algorithm = PG3G_ALG_PATH_TRACING;
    p.min_path_length = 1;
    p.max_path_length = maxPL;
    p.indirect_clipping = 0.f;
    p.splitting_budget = 1.f;
    p.splitting_level = 1.f;
    p.light_to_bsdf_ratio = 1.f;
    p.seed = seed;
    p.scene_id = sceneId;
}
void SetCamera(pg3g_scene_build &b, const float pos[3], const float fwd[3], const float up[3], float hfov, uint32_t rx,
               uint32_t ry) {
    pg3g_camera_desc &c = b.desc.camera;
    for (int k = 0; k < 3; ++k) { c.position[k] = pos[k]; c.forward[k] = fwd[k]; c.up[k] = up[k]; }
    c.hfov_deg = hfov;
    c.res_x = rx;
    c.res_y = ry;
}

// Unit box [0,1]^3 with inward-facing walls and an open front (y = 0).  Materials appended:
// white (floor, ceiling, back), red (left), green (right).
void UnitBox(pg3g_scene_build &b, bool walls) {
    const int white = (int)b.mats.size();
    b.mats.push_back(Lambert(0.8f, 0.8f, 0.8f));
    const int red = (int)b.mats.size();
    b.mats.push_back(Lambert(0.8f, 0.1f, 0.1f));
    const int green = (int)b.mats.size();
    b.mats.push_back(Lambert(0.1f, 0.8f, 0.1f));
    const float P[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    AddQuad(b, P[0], P[1], P[2], P[3], white);             // floor, normal +z
    if (!walls) return;
    AddQuad(b, P[4], P[7], P[6], P[5], white);             // ceiling, normal -z
    AddQuad(b, P[3], P[2], P[6], P[7], white);             // back wall y = 1, normal -y
    AddQuad(b, P[0], P[3], P[7], P[4], red);               // left wall x = 0, normal +x
    AddQuad(b, P[1], P[5], P[6], P[2], green);             // right wall x = 1, normal -x
}
// 0.25 x 0.25 ceiling light of radiance `radiance`, facing down, two emitting triangles.
void CeilingLight(pg3g_scene_build &b, float radiance) {
    const float z = 0.995f, a = 0.375f, c = 0.625f;
    const float q0[3] = {a, a, z}, q1[3] = {a, c, z}, q2[3] = {c, c, z}, q3[3] = {c, a, z};
    const float area = 0.25f * 0.25f / 2.f;
    const float power = radiance * kPi * area;        // AreaLight::SetPower: L = power / (pi * A)
    for (int t = 0; t < 2; ++t) {
        const int lightId = (int)b.lights.size();
        pg3g_material m = Lambert(0.f, 0.f, 0.f);
        m.light_id = lightId;
        const int mid = (int)b.mats.size();
        b.mats.push_back(m);
        if (t == 0) { AddTri(b, q0, q1, q2, mid); b.lights.push_back(AreaLight(q0, q1, q2, power)); }
        else { AddTri(b, q0, q2, q3, mid); b.lights.push_back(AreaLight(q0, q2, q3, power)); }
    }
}
void BoxCamera(pg3g_scene_build &b, uint32_t rx, uint32_t ry) {
    const float pos[3] = {0.5f, -1.25f, 0.5f}, fwd[3] = {0.f, 1.f, 0.f}, up[3] = {0.f, 0.f, 1.f};
    SetCamera(b, pos, fwd, up, 45.f, rx, ry);
}

// Synthetic lat-long EM (texel centres mapped with LatLong2Dir, geom.hxx:65-76): a sky gradient of
// luminance lo..hi from nadir to zenith, plus Gaussian lobes (direction, sigma in radians, peak).
struct Lobe { float d[3]; float sigma; float peak; float rgb[3]; };
void SynthEm(pg3g_scene_build &b, uint32_t w, uint32_t h, float lo, float hi, const std::vector<Lobe> &lobes) {
    b.em.assign((size_t)w * h * 3, 0.f);
    const float sky[3] = {0.5f, 0.75f, 1.0f};
    const float skyLum = 0.212671f * sky[0] + 0.715160f * sky[1] + 0.072169f * sky[2];
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            const double u = (x + 0.5) / w, v = (y + 0.5) / h;
            const double phi = -(u - 0.5) * 2 * 3.14159265358979323846, th = v * 3.14159265358979323846;
            const double d[3] = {std::sin(th) * std::cos(phi), std::sin(th) * std::sin(phi), std::cos(th)};
            const double lum = lo + (hi - lo) * (0.5 + 0.5 * d[2]);
            float *t = &b.em[((size_t)y * w + x) * 3];
            for (int c = 0; c < 3; ++c) t[c] = (float)(sky[c] / skyLum * lum);
            for (const Lobe &l : lobes) {
                const double cs = std::max(-1.0, std::min(1.0, d[0] * l.d[0] + d[1] * l.d[1] + d[2] * l.d[2]));
                const double ang = std::acos(cs);
                const double g = l.peak * std::exp(-ang * ang / (2.0 * l.sigma * l.sigma));
                for (int c = 0; c < 3; ++c) t[c] += (float)(g * l.rgb[c]);
            }
        }
    b.desc.em.width = w;
    b.desc.em.height = h;
    b.desc.em.bilinear = 1;
}
// Sun direction that lights the box interior through its open front: from -y, above the horizon.
void SunDir(Rand &r, float d[3]) {
    const double theta = r.U(45.0, 75.0) * 3.14159265358979323846 / 180.0;
    const double phi = (-90.0 + r.U(-40.0, 40.0)) * 3.14159265358979323846 / 180.0;
    d[0] = (float)(std::sin(theta) * std::cos(phi));
    d[1] = (float)(std::sin(theta) * std::sin(phi));
    d[2] = (float)std::cos(theta);
}

int BuildWorkload(pg3g_scene_build &b, int workload, uint64_t seed, uint32_t variant, uint32_t res) {
    Rand r{seed * 0x2545F4914F6CDD1Dull + 0x1234567ull + (uint64_t)variant * 0x9E3779B97F4A7C15ull};
    switch (workload) {
    case 1: {
        UnitBox(b, true);
        CeilingLight(b, 25.f);
        const int diffuse = (int)b.mats.size();
        b.mats.push_back(Lambert(0.6f, 0.6f, 0.6f));
        const int metal = (int)b.mats.size();
        b.mats.push_back(GgxConductor(0.1f, 0.95f));
        AddSphere(b, 0.3f, 0.6f, 0.25f, 0.25f, diffuse);
        AddSphere(b, 0.7f, 0.35f, 0.25f, 0.25f, metal);
        const uint32_t n = res ? res : 256;
        BoxCamera(b, n, n);
        SetParams(b, 8, seed, 0);
        b.spp = 16;
        return PG3G_OK;
    }
    case 2: {
        UnitBox(b, true);
        const int diffuse = (int)b.mats.size();
        b.mats.push_back(Lambert(0.6f, 0.6f, 0.6f));
        const int glass = (int)b.mats.size();
        b.mats.push_back(GgxDielectric(0.3f, 1.5f));
        AddSphere(b, 0.3f, 0.6f, 0.25f, 0.25f, diffuse);
        AddSphere(b, 0.7f, 0.35f, 0.25f, 0.25f, glass);
        Lobe sun;
        SunDir(r, sun.d);
        sun.sigma = 2.0f * kPi / 180.f;
        sun.peak = 1e4f;
        sun.rgb[0] = sun.rgb[1] = sun.rgb[2] = 1.f;
        SynthEm(b, 1024, 512, 0.2f, 1.0f, {sun});
        b.lights.push_back(Background(0.f, 0.f, 0.f));
        b.desc.em.sampler = PG3G_EM_STEERABLE;
        b.desc.em.st_max_approx_error = 5.f;
        b.desc.em.st_min_subdiv_level = 4;
        b.desc.em.st_max_subdiv_level = 4;
        const uint32_t n = res ? res : 512;
        BoxCamera(b, n, n);
        SetParams(b, 10, seed, 0);
        b.spp = 64;
        return PG3G_OK;
    }
    case 3: {
        if (variant >= 64) return PG3G_ERR_INVALID_ARG;
        UnitBox(b, true);
        CeilingLight(b, 25.f);
        const float alb = (float)r.U(0.1, 0.9);
        const float alpha = (float)std::exp(r.U(std::log(0.01), std::log(0.5)));
        const float ior = (float)r.U(1.3, 1.8);
        const int diffuse = (int)b.mats.size();
        b.mats.push_back(Lambert(alb, alb, alb));
        const int glass = (int)b.mats.size();
        b.mats.push_back(GgxDielectric(alpha, ior));
        AddSphere(b, 0.3f, 0.6f, 0.25f, 0.25f, diffuse);
        AddSphere(b, 0.7f, 0.35f, 0.25f, 0.25f, glass);
        Lobe sun;
        SunDir(r, sun.d);
        sun.sigma = 2.0f * kPi / 180.f;
        sun.peak = (float)std::exp(r.U(std::log(1e2), std::log(1e4)));
        sun.rgb[0] = sun.rgb[1] = sun.rgb[2] = 1.f;
        SynthEm(b, 256, 128, 0.2f, 1.0f, {sun});
        b.lights.push_back(Background(0.f, 0.f, 0.f));
        b.desc.em.sampler = PG3G_EM_STEERABLE;
        b.desc.em.st_max_approx_error = 5.f;
        b.desc.em.st_min_subdiv_level = 4;
        b.desc.em.st_max_subdiv_level = 4;
        const uint32_t side = res ? res : (uint32_t)std::lround(r.U(64.0, 512.0));
        BoxCamera(b, side, side);
        const uint32_t spps[4] = {4, 16, 64, 256};
        b.spp = spps[r.Next() % 4];
        SetParams(b, 8, seed, variant);
        return PG3G_OK;
    }
    case 4: {
        UnitBox(b, true);
        const int a = (int)b.mats.size();
        pg3g_material m0 = GgxConductor((float)r.U(0.05, 0.5), 0.95f);
        pg3g_material m1 = GgxConductor((float)r.U(0.05, 0.5), 0.95f);
        for (int c = 0; c < 3; ++c) { m0.glossy[c] = (float)r.U(0.2, 0.9); m1.glossy[c] = (float)r.U(0.2, 0.9); }
        b.mats.push_back(m0);
        b.mats.push_back(m1);
        AddSphere(b, 0.3f, 0.6f, 0.25f, 0.25f, a);
        AddSphere(b, 0.7f, 0.35f, 0.25f, 0.25f, a + 1);
        std::vector<Lobe> lobes;
        for (int i = 0; i < 2; ++i) {
            Lobe l;
            SunDir(r, l.d);
            l.sigma = (float)r.U(5.0, 15.0) * kPi / 180.f;
            l.peak = (float)r.U(5.0, 40.0);
            for (int c = 0; c < 3; ++c) l.rgb[c] = (float)r.U(0.5, 1.0);
            lobes.push_back(l);
        }
        SynthEm(b, 128, 64, 0.2f, 1.0f, lobes);
        b.lights.push_back(Background(0.f, 0.f, 0.f));
        b.desc.em.sampler = PG3G_EM_SIMPLE;
        const uint32_t n = res ? res : 256;
        BoxCamera(b, n, n);
        SetParams(b, 8, seed, 0);
        b.spp = 32;
        return PG3G_OK;
    }
    case 5: {
        if (variant >= 8) return PG3G_ERR_INVALID_ARG;
        // one scene seen from 8 views: materials and EM depend on the seed only, the variant
        // picks the camera
        r = Rand{seed * 0x2545F4914F6CDD1Dull + 0x1234567ull};
        UnitBox(b, false);   // floor only: the 8 views circle the spheres
        const int a = (int)b.mats.size();
        const float xs[4] = {0.3f, 0.7f, 0.3f, 0.7f}, ys[4] = {0.3f, 0.3f, 0.7f, 0.7f};
        for (int i = 0; i < 4; ++i) {
            b.mats.push_back(GgxConductor((float)r.U(0.05, 0.5), 0.95f));
            AddSphere(b, xs[i], ys[i], 0.15f, 0.15f, a + i);
        }
        std::vector<Lobe> lobes;
        for (int i = 0; i < 3; ++i) {
            Lobe l;
            const double th = r.U(10.0, 80.0) * 3.14159265358979323846 / 180.0, ph = r.U(0.0, 2 * 3.14159265358979323846);
            l.d[0] = (float)(std::sin(th) * std::cos(ph)); l.d[1] = (float)(std::sin(th) * std::sin(ph)); l.d[2] = (float)std::cos(th);
            l.sigma = (float)r.U(5.0, 15.0) * kPi / 180.f;
            l.peak = (float)r.U(5.0, 40.0);
            for (int c = 0; c < 3; ++c) l.rgb[c] = (float)r.U(0.5, 1.0);
            lobes.push_back(l);
        }
        SynthEm(b, 128, 64, 0.2f, 1.0f, lobes);
        b.lights.push_back(Background(0.f, 0.f, 0.f));
        b.desc.em.sampler = PG3G_EM_SIMPLE;
        const double ang = variant * 2.0 * 3.14159265358979323846 / 8.0;
        const float pos[3] = {(float)(0.5 + 1.6 * std::sin(ang)), (float)(0.5 - 1.6 * std::cos(ang)), 0.9f};
        const float tgt[3] = {0.5f, 0.5f, 0.15f};
        const float fwd[3] = {tgt[0] - pos[0], tgt[1] - pos[1], tgt[2] - pos[2]};
        const float up[3] = {0.f, 0.f, 1.f};
        const uint32_t n = res ? res : 256;
        SetCamera(b, pos, fwd, up, 45.f, n, n);
        SetParams(b, 6, seed, variant);
        b.spp = 16;
        return PG3G_OK;
    }
    default:
        return PG3G_ERR_INVALID_ARG;
    }
}

// --- reference scenes (scene.hxx:176-956) ---------------------------------------------------------
enum : uint32_t {
    kLightCeiling = 0x1, kLightBox = 0x2, kLightPoint = 0x4, kLightEnv = 0x8,
    k2Spheres = 0x10, k1Sphere = 0x20, kVerticalRectangle = 0x40, kDiagonalRectangles = 0x80,
    kWalls = 0x100, kFloor = 0x200, kLayeredSphere = 0x10000000,
    kSpheresPhongDiffuse = 0x1000, kSpheresPhongGlossy = 0x2000, kWallsPhongDiffuse = 0x4000,
    kWallsPhongGlossy = 0x8000, kSpheresFresnelConductor = 0x10000, kSpheresFresnelDielectric = 0x20000,
    kSpheresMicrofacetGGXConductor = 0x40000, kSpheresMicrofacetGGXDielectric = 0x80000,
    kSpheresSraLayers = 0x100000, kVertRectFresnelDielectric = 0x200000, kVertRectMicrofacetGGXDielectric = 0x400000,
};
const uint32_t kFullBox = kWalls | kFloor | k2Spheres, k2SpheresOnFloor = kFloor | k2Spheres,
               kBox1Sphere = kWalls | kFloor | k1Sphere, k1SphereOnly = k1Sphere,
               kRects = kVerticalRectangle | kDiagonalRectangles,
               kMatsDiffuse = kWallsPhongDiffuse | kSpheresPhongDiffuse, kMatsGlossy = kWallsPhongGlossy | kSpheresPhongGlossy;
// g_SceneConfigs, config.hxx:45-115
const uint32_t kSceneConfigs[] = {
    kLightPoint | kFullBox | kMatsDiffuse, kLightPoint | kFullBox | kMatsDiffuse | kMatsGlossy,
    kLightCeiling | kFullBox | kMatsDiffuse, kLightCeiling | kFullBox | kMatsDiffuse | kMatsGlossy,
    kLightBox | kFullBox | kMatsDiffuse, kLightBox | kFullBox | kMatsDiffuse | kMatsGlossy,
    kLightEnv | kFullBox | kMatsDiffuse, kLightEnv | kFullBox | kMatsDiffuse | kMatsGlossy,
    kLightEnv | k2SpheresOnFloor | kMatsDiffuse, kLightEnv | k2SpheresOnFloor | kMatsDiffuse | kMatsGlossy,
    kLightBox | kLightEnv | kFullBox | kMatsDiffuse, kLightPoint | kLightBox | kFullBox | kMatsDiffuse,
    kLightPoint | kLightBox | kLightEnv | kFullBox | kMatsDiffuse,
    kLightPoint | kLightBox | kLightEnv | k2SpheresOnFloor | kMatsDiffuse,
    kLightPoint | kLightBox | kLightEnv | k2SpheresOnFloor | kMatsDiffuse | kMatsGlossy,
    kLightPoint | kLightBox | kLightEnv | kFullBox | kMatsDiffuse | kMatsGlossy,
    kLightPoint | kFullBox | kSpheresFresnelConductor | kWallsPhongDiffuse,
    kLightBox | kFullBox | kSpheresFresnelConductor | kWallsPhongDiffuse,
    kLightCeiling | kFullBox | kSpheresFresnelConductor | kWallsPhongDiffuse,
    kLightBox | kFullBox | kSpheresFresnelConductor | kWallsPhongDiffuse | kWallsPhongGlossy,
    kLightEnv | k1SphereOnly | kSpheresPhongDiffuse, kLightEnv | k1SphereOnly | kSpheresPhongGlossy,
    kLightEnv | k1SphereOnly | kSpheresPhongDiffuse | kSpheresPhongGlossy,
    kLightEnv | k1SphereOnly | kSpheresFresnelConductor, kLightEnv | k1SphereOnly | kSpheresFresnelDielectric,
    kLightEnv | k1SphereOnly | kSpheresMicrofacetGGXConductor, kLightEnv | k1SphereOnly | kSpheresMicrofacetGGXDielectric,
    kLightEnv | k1SphereOnly | kSpheresSraLayers,
    kLightEnv | kRects | kVertRectFresnelDielectric, kLightEnv | kRects | kVertRectMicrofacetGGXDielectric,
    kLightBox | kBox1Sphere | kSpheresFresnelConductor | kWallsPhongDiffuse,
    kLightBox | kBox1Sphere | kSpheresFresnelDielectric | kWallsPhongDiffuse,
    kLightBox | kBox1Sphere | kSpheresMicrofacetGGXConductor | kWallsPhongDiffuse,
    kLightBox | kBox1Sphere | kSpheresMicrofacetGGXDielectric | kWallsPhongDiffuse,
    kLightCeiling | kBox1Sphere | kSpheresMicrofacetGGXConductor | kWallsPhongDiffuse,
    kLightCeiling | kBox1Sphere | kSpheresMicrofacetGGXDielectric | kWallsPhongDiffuse,
    kLightPoint | kBox1Sphere | kSpheresMicrofacetGGXDielectric | kWallsPhongDiffuse,
    kLightEnv | kBox1Sphere | kSpheresMicrofacetGGXDielectric | kWallsPhongDiffuse,
};
// SpectralData, materials.hxx:19-32
const float kAirIor = 1.000277f, kGlassIor = 1.510000f, kCopperIor = 0.468000f, kCopperAbsorb = 2.810000f,
            kSilverIor = 0.121000f, kSilverAbsorb = 3.660000f;

pg3g_material Phong(float r, float g, float bl, float glossy, float exponent, bool dOn, bool gOn) {
    pg3g_material m = Mat(PG3G_MAT_PHONG);
    m.diffuse[0] = r; m.diffuse[1] = g; m.diffuse[2] = bl;
    m.glossy[0] = m.glossy[1] = m.glossy[2] = glossy;
    m.phong_exponent = exponent;
    m.phong_diffuse_on = dOn ? 1 : 0;
    m.phong_glossy_on = gOn ? 1 : 0;
    return m;
}
pg3g_material Conductor(int type, float alpha, float inner, float outer, float k) {
    pg3g_material m = Mat(type);
    m.alpha = alpha; m.inner_ior = inner; m.outer_ior = outer; m.absorbance = k;
    return m;
}
pg3g_material Dielectric(int type, float alpha, float inner, float outer, bool back) {
    pg3g_material m = Mat(type);
    m.alpha = alpha; m.inner_ior = inner; m.outer_ior = outer; m.back_side_light_sampling = back ? 1 : 0;
    return m;
}

// Image EMs of the reference scenes (scene.hxx:850-943): azimuth rotation and scale given to
// LoadEnvironmentMap for em types 2..12.  Types 2 and 3 are constant white images (1.0 in every
// texel, 8x4 and 1024x512) and are built in memory; the others are photographs loaded from a file.
struct RefEm { float rotation, scale; uint32_t w, h; };
const RefEm kRefEms[13] = {{0, 0, 0, 0},       {0, 0, 0, 0},      {0.f, 1.f, 8, 4},   {0.f, 1.f, 1024, 512},
                           {0.40f, 10.f, 0, 0}, {0.05f, 1.f, 0, 0}, {0.05f, 1.f, 0, 0}, {0.83f, 1.5f, 0, 0},
                           {0.1f, 2.f, 0, 0},  {0.53f, 0.2f, 0, 0}, {-0.12f, 0.5f, 0, 0}, {-0.12f, 3.f, 0, 0},
                           {0.f, 7500.f, 0, 0}};

int BuildReference(pg3g_scene_build &b, uint32_t id, uint32_t em, uint32_t res, const char *emPath = nullptr) {
    if (id >= sizeof(kSceneConfigs) / sizeof(kSceneConfigs[0])) return PG3G_ERR_INVALID_ARG;
    const uint32_t M = kSceneConfigs[id];
    if (M & (kSpheresSraLayers | kLayeredSphere)) return PG3G_ERR_UNSUPPORTED;
    const bool ceiling = M & kLightCeiling, box = M & kLightBox, point = M & kLightPoint, env = M & kLightEnv;
    if (env && em > 12) return PG3G_ERR_INVALID_ARG;
    if (env && em > 3 && !emPath) return PG3G_ERR_UNSUPPORTED;   // photographs: the caller names the file
    const uint32_t n = res ? res : 512;
    const float pos[3] = {-0.0439815f, -4.12529f, 0.222539f}, fwd[3] = {0.00688625f, 0.998505f, -0.0542161f},
                up[3] = {3.73896e-4f, 0.0542148f, 0.998529f};
    SetCamera(b, pos, fwd, up, 45.f, n, n);
    // materials 0..12 (scene.hxx:203-434)
    b.mats.push_back(Lambert(0, 0, 0));
    b.mats.push_back(Lambert(0, 0, 0));
    const bool wd = M & kWallsPhongDiffuse, wg = M & kWallsPhongGlossy, sd = M & kSpheresPhongDiffuse, sg = M & kSpheresPhongGlossy;
    b.mats.push_back(Phong(0.803922f, 0.803922f, 0.803922f, 0.5f, 90, wd, wg));
    b.mats.push_back(Phong(0.156863f, 0.803922f, 0.172549f, 0.5f, 90, wd, wg));
    b.mats.push_back(Phong(0.803922f, 0.152941f, 0.152941f, 0.5f, 90, wd, wg));
    b.mats.push_back(Phong(0.803922f, 0.803922f, 0.803922f, 0.5f, 90, wd, wg));
    if (M & kSpheresFresnelConductor) {
        b.mats.push_back(Conductor(PG3G_MAT_SMOOTH_CONDUCTOR, 0.f, kCopperIor, kAirIor, kCopperAbsorb));
        b.mats.push_back(Conductor(PG3G_MAT_SMOOTH_CONDUCTOR, 0.f, kSilverIor, kAirIor, kSilverAbsorb));
    } else if (M & kSpheresMicrofacetGGXConductor) {
        b.mats.push_back(Conductor(PG3G_MAT_GGX_CONDUCTOR, 0.1f, kCopperIor, kAirIor, kCopperAbsorb));
        b.mats.push_back(Conductor(PG3G_MAT_GGX_CONDUCTOR, 0.1f, kCopperIor, kAirIor, kCopperAbsorb));
    } else {
        b.mats.push_back(Phong(0.803922f, 0.803922f, 0.152941f, 0.7f, 200, sd, sg));
        b.mats.push_back(Phong(0.152941f, 0.152941f, 0.803922f, 0.7f, 600, sd, sg));
    }
    if (M & kSpheresFresnelConductor) b.mats.push_back(Conductor(PG3G_MAT_SMOOTH_CONDUCTOR, 0.f, kCopperIor, kAirIor, kCopperAbsorb));
    else if (M & kSpheresFresnelDielectric) b.mats.push_back(Dielectric(PG3G_MAT_SMOOTH_DIELECTRIC, 0.f, kGlassIor, kAirIor, false));
    else if (M & kSpheresMicrofacetGGXConductor) b.mats.push_back(Conductor(PG3G_MAT_GGX_CONDUCTOR, 0.01f, kCopperIor, kAirIor, kCopperAbsorb));
    else if (M & kSpheresMicrofacetGGXDielectric) b.mats.push_back(Dielectric(PG3G_MAT_GGX_DIELECTRIC, 0.01f, kGlassIor, kAirIor, false));
    else b.mats.push_back(Phong(0.803922f, 0.803922f, 0.803922f, 0.5f, 90, sd, sg));
    if (M & kVertRectFresnelDielectric) b.mats.push_back(Dielectric(PG3G_MAT_SMOOTH_DIELECTRIC, 0.f, kGlassIor, kAirIor, false));
    else if (M & kVertRectMicrofacetGGXDielectric) b.mats.push_back(Dielectric(PG3G_MAT_GGX_DIELECTRIC, 0.1f, kGlassIor, kAirIor, true));
    else b.mats.push_back(Phong(0.8f, 0.8f, 0.8f, 0.0f, 1, true, false));
    b.mats.push_back(Phong(0.8f, 0.8f, 0.8f, 0.0f, 1, true, false));

    const float cb[8][3] = {{-1.27029f, 1.30455f, -1.28002f}, {1.28975f, 1.30455f, -1.28002f}, {1.28975f, 1.30455f, 1.28002f},
                            {-1.27029f, 1.30455f, 1.28002f}, {-1.27029f, -1.25549f, -1.28002f}, {1.28975f, -1.25549f, -1.28002f},
                            {1.28975f, -1.25549f, 1.28002f}, {-1.27029f, -1.25549f, 1.28002f}};
    if (M & kFloor) { AddTri(b, cb[0], cb[4], cb[5], 2); AddTri(b, cb[5], cb[1], cb[0], 2); }
    if (M & kWalls) {
        AddTri(b, cb[3], cb[7], cb[4], 3); AddTri(b, cb[4], cb[0], cb[3], 3);
        AddTri(b, cb[1], cb[5], cb[6], 4); AddTri(b, cb[6], cb[2], cb[1], 4);
        AddTri(b, cb[0], cb[1], cb[2], 5); AddTri(b, cb[2], cb[3], cb[0], 5);
        if (ceiling && !box) { AddTri(b, cb[2], cb[6], cb[7], 0); AddTri(b, cb[7], cb[3], cb[2], 1); }
        else { AddTri(b, cb[2], cb[6], cb[7], 2); AddTri(b, cb[7], cb[3], cb[2], 2); }
    }
    if (M & k2Spheres) {
        const float r = 0.5f;
        const float lw[3] = {(cb[0][0] + cb[4][0]) * 0.5f, (cb[0][1] + cb[4][1]) * 0.5f, (cb[0][2] + cb[4][2]) * 0.5f + r};
        const float rw[3] = {(cb[1][0] + cb[5][0]) * 0.5f, (cb[1][1] + cb[5][1]) * 0.5f, (cb[1][2] + cb[5][2]) * 0.5f + r};
        const float w = rw[0] - lw[0];
        AddSphere(b, lw[0] + 2.f * w / 7.f, lw[1], lw[2], r, 6);
        AddSphere(b, rw[0] - 2.f * w / 7.f, rw[1] + w / 4, rw[2], r, 7);
    }
    if (M & k1Sphere) {
        const float r = 1.f;
        AddSphere(b, (cb[0][0] + cb[5][0]) * 0.5f, (cb[0][1] + cb[5][1]) * 0.5f, (cb[0][2] + cb[5][2]) * 0.5f + r, r, 8);
    }
    auto add = [](const float *a, float x, float y, float z, float *o) { o[0] = a[0] + x; o[1] = a[1] + y; o[2] = a[2] + z; };
    if (M & kDiagonalRectangles) {
        float fc[3], cc[3];
        for (int k = 0; k < 3; ++k) { fc[k] = (cb[4][k] + cb[5][k]) * 0.5f; cc[k] = (cb[3][k] + cb[2][k]) * 0.5f; }
        const float hgt = std::fabs(cb[0][2] - cb[3][2]), fw = std::fabs(cb[0][0] - cb[1][0]);
        const float hw = 0.15f * fw * 0.5f, xo = 0.6f * fw * 0.5f, zo = 0.1f * hgt;
        for (int s = -1; s <= 1; s += 2) {
            float q[4][3];
            add(fc, -hw + s * xo, 0, zo, q[0]); add(fc, hw + s * xo, 0, zo, q[1]);
            add(cc, hw + s * xo, 0, zo, q[2]); add(cc, -hw + s * xo, 0, zo, q[3]);
            AddTri(b, q[3], q[0], q[1], 10); AddTri(b, q[1], q[2], q[3], 10);
        }
    }
    if (M & kVerticalRectangle) {
        float fc[3];
        for (int k = 0; k < 3; ++k) fc[k] = (cb[0][k] + cb[5][k]) * 0.5f;
        const float fw = std::fabs(cb[0][0] - cb[1][0]), hgt = std::fabs(cb[0][2] - cb[3][2]), dep = std::fabs(cb[0][1] - cb[4][1]);
        const float hw = 8.0f * fw * 0.5f, rh = 0.7f * hgt, yo = -1.0f * dep * 0.5f, zo = 0.1f * hgt;
        float q[4][3];
        add(fc, -hw, yo, zo, q[0]); add(fc, hw, yo, zo, q[1]); add(fc, hw, yo, zo + rh, q[2]); add(fc, -hw, yo, zo + rh, q[3]);
        AddTri(b, q[3], q[0], q[1], 9); AddTri(b, q[1], q[2], q[3], 9);
    }
    const float lb[8][3] = {{-0.25f, 0.25f, 1.26002f}, {0.25f, 0.25f, 1.26002f}, {0.25f, 0.25f, 1.28002f}, {-0.25f, 0.25f, 1.28002f},
                            {-0.25f, -0.25f, 1.26002f}, {0.25f, -0.25f, 1.26002f}, {0.25f, -0.25f, 1.28002f}, {-0.25f, -0.25f, 1.28002f}};
    if (box && !ceiling) {
        AddTri(b, lb[0], lb[2], lb[1], 5); AddTri(b, lb[2], lb[0], lb[3], 5);
        AddTri(b, lb[3], lb[4], lb[7], 5); AddTri(b, lb[4], lb[3], lb[0], 5);
        AddTri(b, lb[1], lb[6], lb[5], 5); AddTri(b, lb[6], lb[1], lb[2], 5);
        AddTri(b, lb[4], lb[5], lb[6], 5); AddTri(b, lb[6], lb[7], lb[4], 5);
        AddTri(b, lb[0], lb[5], lb[4], 0); AddTri(b, lb[5], lb[0], lb[1], 1);
    }
    // lights (scene.hxx:767-955): 25 W split over two triangles
    if (ceiling && !box) {
        b.lights.push_back(AreaLight(cb[2], cb[6], cb[7], 12.5f));
        b.lights.push_back(AreaLight(cb[7], cb[3], cb[2], 12.5f));
        b.mats[0].light_id = 0;
        b.mats[1].light_id = 1;
    }
    if (box && !ceiling) {
        b.lights.push_back(AreaLight(lb[0], lb[5], lb[4], 12.5f));
        b.lights.push_back(AreaLight(lb[5], lb[0], lb[1], 12.5f));
        b.mats[0].light_id = 0;
        b.mats[1].light_id = 1;
    }
    if (point) {
        pg3g_light l;
        std::memset(&l, 0, sizeof(l));
        l.type = PG3G_LIGHT_POINT;
        l.p0[0] = 0.0f; l.p0[1] = -0.5f; l.p0[2] = 1.0f;
        l.power[0] = l.power[1] = l.power[2] = 50.f;
        b.lights.push_back(l);
    }
    if (env) {
        if (em == 0) b.lights.push_back(Background(135 / 255.f, 206 / 255.f, 250 / 255.f));
        else if (em == 1) b.lights.push_back(Background(0.87f, 0.87f, 0.87f));
        else {
            const RefEm &r = kRefEms[em];
            uint32_t w = r.w, h = r.h;
            if (emPath) {
                std::string err;
                if (!pg3::LoadEmFile(emPath, r.rotation, r.scale, b.em, w, h, err)) return PG3G_ERR_INVALID_ARG;
            } else {
                b.em.assign((size_t)w * h * 3, 1.f * r.scale);
            }
            b.lights.push_back(Background(0.f, 0.f, 0.f));
            b.desc.em.width = w;
            b.desc.em.height = h;
            b.desc.em.bilinear = 1;                     // LoadEnvironmentMap default (lights.hxx:372)
            b.desc.em.sampler = PG3G_EM_SIMPLE;         // the reference's default sampler (hard_config.hxx:17)
        }
    }
    SetParams(b, 0, 1234, 0);
    b.desc.params.indirect_clipping = 16.f;   // Config defaults, config.hxx:567-578
    b.desc.params.splitting_budget = 4.f;
    b.spp = 1;
    return PG3G_OK;
}

}  // namespace

extern "C" {

int pg3g_build_workload_scene(int workload, uint64_t seed, uint32_t variant, uint32_t res_override, pg3g_scene_build **out) {
    if (!out) return PG3G_ERR_INVALID_ARG;
    *out = nullptr;
    pg3g_scene_build *b = new (std::nothrow) pg3g_scene_build();
    if (!b) return PG3G_ERR_OUT_OF_MEMORY;
    const int rc = BuildWorkload(*b, workload, seed, variant, res_override);
    if (rc != PG3G_OK) { delete b; return rc; }
    b->Finish();
    *out = b;
    return PG3G_OK;
}

int pg3g_reference_scene_acronym(uint32_t scene_id, uint32_t em_type, char *buf, size_t len) {
    if (!buf || len == 0 || scene_id >= sizeof(kSceneConfigs) / sizeof(kSceneConfigs[0])) return PG3G_ERR_INVALID_ARG;
    const uint32_t M = kSceneConfigs[scene_id];
    std::string a;
    // geometry
    const size_t g0 = a.size();
    if (M & kWalls) a += "w";
    if (M & k2Spheres) a += "2s";
    if (M & k1Sphere) a += "1s";
    if (M & kLayeredSphere) a += "ls";
    if (M & kVerticalRectangle) a += "vr";
    if (M & kDiagonalRectangles) a += "dr";
    if (a.size() == g0 && !(M & (kWalls | kFloor | k2Spheres | k1Sphere | kLayeredSphere | kRects))) a += "e";
    a += "_";
    // lights (the environment light carries its EM type number)
    if (M & kLightCeiling) a += "c";
    if (M & kLightBox) a += "b";
    if (M & kLightPoint) a += "p";
    if (M & kLightEnv) a += "e" + std::to_string(em_type);
    a += "_";
    // materials
    if (M & (kSpheresPhongDiffuse | kSpheresPhongGlossy)) {
        a += "Sp";
        if (M & kSpheresPhongDiffuse) a += "d";
        if (M & kSpheresPhongGlossy) a += "g";
    } else if (M & kSpheresFresnelConductor) a += "Sffc";
    else if (M & kSpheresFresnelDielectric) a += "Sffd";
    else if (M & kSpheresMicrofacetGGXConductor) a += "Smgc";
    else if (M & kSpheresMicrofacetGGXDielectric) a += "Smgd";
    else if (M & kSpheresSraLayers) a += "Ssral";
    if (M & (kWallsPhongDiffuse | kWallsPhongGlossy)) {
        a += "Wp";
        if (M & kWallsPhongDiffuse) a += "d";
        if (M & kWallsPhongGlossy) a += "g";
    }
    if (M & kVerticalRectangle) {
        if (M & kVertRectFresnelDielectric) a += "Rffd";
        else if (M & kVertRectMicrofacetGGXDielectric) a += "Rmgd";
        else a += "Rpd";
    }
    if (a.size() + 1 > len) return PG3G_ERR_INVALID_ARG;
    std::memcpy(buf, a.c_str(), a.size() + 1);
    return PG3G_OK;
}

int pg3g_build_reference_scene(uint32_t scene_id, uint32_t em_type, uint32_t resolution, pg3g_scene_build **out) {
    if (!out) return PG3G_ERR_INVALID_ARG;
    *out = nullptr;
    pg3g_scene_build *b = new (std::nothrow) pg3g_scene_build();
    if (!b) return PG3G_ERR_OUT_OF_MEMORY;
    const int rc = BuildReference(*b, scene_id, em_type, resolution);
    if (rc != PG3G_OK) { delete b; return rc; }
    b->Finish();
    *out = b;
    return PG3G_OK;
}

int pg3g_build_reference_scene_em(uint32_t scene_id, uint32_t em_type, uint32_t resolution, const char *em_path,
                                  pg3g_scene_build **out) {
    if (!out) return PG3G_ERR_INVALID_ARG;
    *out = nullptr;
    pg3g_scene_build *b = new (std::nothrow) pg3g_scene_build();
    if (!b) return PG3G_ERR_OUT_OF_MEMORY;
    const int rc = BuildReference(*b, scene_id, em_type, resolution, em_path);
    if (rc != PG3G_OK) { delete b; return rc; }
    b->Finish();
    *out = b;
    return PG3G_OK;
}

const pg3g_scene_desc *pg3g_scene_build_desc(const pg3g_scene_build *b) { return b ? &b->desc : nullptr; }
pg3g_scene_desc *pg3g_scene_build_mutable(pg3g_scene_build *b) { return b ? &b->desc : nullptr; }
float *pg3g_scene_build_em_texels(pg3g_scene_build *b) { return (b && !b->em.empty()) ? b->em.data() : nullptr; }
uint32_t pg3g_scene_build_spp(const pg3g_scene_build *b) { return b ? b->spp : 0; }
void pg3g_scene_build_free(pg3g_scene_build *b) { delete b; }

}  // extern "C"
