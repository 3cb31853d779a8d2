// C ABI of libpg3g: contexts, scene upload with EM table builds, renders, framebuffer access and
// the function-level entry points.  No exception crosses the ABI; failures return a status and
// leave a message for pg3g_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "pg3g.h"
#include "pg3g_internal.hpp"

using namespace pg3;

namespace {

thread_local std::string g_lastError;

int Fail(int code, const std::string &msg) {
    g_lastError = msg;
    return code;
}

#define PG3_HIP(expr)                                                                         \
    do {                                                                                      \
        const hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                                 \
            return Fail(e_ == hipErrorOutOfMemory ? PG3G_ERR_OUT_OF_MEMORY : PG3G_ERR_HIP,    \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                   \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { Free(); }
    void Free() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t Alloc(size_t n) {
        if (n <= bytes && p) return hipSuccess;
        Free();
        if (n == 0) return hipSuccess;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

struct SceneSlot {
    bool valid = false;
    SceneView view{};
    pg3g_render_params params{};
    DevBuf quads, tris, sphs, mats, lights, texels, lum, margCdf, condCdf, rowInt, nodeW, leafVerts, fb;
    DevBuf viewDev;      // device copy of `view`, refreshed on upload and parameter changes
    // differentiable rendering
    std::vector<pg3g_material> hostMats;      // sampling materials as given (ctor arguments)
    std::vector<pg3g_material> evalHostMats;  // evaluation materials
    std::vector<float> hostTexels;            // sampling EM texels (w*h*3) for table rebuilds
    std::vector<float> evalHostTexels;        // evaluation EM texels when they differ (evalTexDirty)
    bool evalTexDirty = false;
    pg3g_em_desc emDesc{};
    DevBuf evalMats, evalTexels, gradBase, gradScal, gradEm, adjImg, adjWork;
    DevBuf sampleL;                           // per-sample radiance of the last differentiable forward
    uint32_t sampleViews = 1;                 // views of that forward
    uint32_t numViews = 0;                    // cameras of pg3g_diff_set_views (0: none)
    DevBuf viewCams, fbViews;                 // DevView per view; numViews FP64 images
    bool sampleValid = false;
    uint32_t sampleFirst = 0, sampleCount = 0;
    uint64_t sampleEpoch = 0;
    uint64_t epoch = 0;                       // bumped by every change of this slot's scene view
    std::vector<uint32_t> diffMats;           // differentiated materials, scalar layout order
    std::vector<int32_t> gradBaseHost;
    uint32_t numScalars = 0;
    uint64_t iterations = 0;
    uint32_t pendFirst = 0, pendCount = 0;    // iterations queued by pg3g_run_iteration, not launched
    SteerableTables steer;
    bool hasSteer = false;
};

}  // namespace

struct pg3g_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<std::unique_ptr<SceneSlot>> slots;
    DevBuf counters;   // [segments, camera samples]
    DevBuf work;       // item counter of the persistent path-tracing kernel
    int kernelMode = PG3G_KERNEL_MEGAKERNEL;   // the fused kernel: 3-4x the wavefront pipeline
    DevBuf scratch[8];
    WavefrontContext *wavefront = nullptr;
    uint64_t sceneEpoch = 0;   // bumped by every upload / parameter change (invalidates batches)
    int numCUs = 0;
    bool noAllocReady = false; // the scratch array of pg3g_adjoint_noalloc was found on this device
};

struct pg3g_batch {
    pg3g_context *ctx = nullptr;
    uint64_t epoch = 0;
    std::vector<uint32_t> slots, spp;
    std::vector<BatchScene> info;
    DevBuf views, infoDev, work;
    uint32_t gridBlocks = 0, blockLanes = 0;
    pg3g_batch_config cfg{8, 16, 1};
    uint32_t numItems = 0;
    uint32_t passes = 0;       // passes rendered since creation / the last autotune
    int emKind = 0, oneLight = -1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~pg3g_batch() { DropGraph(); }
    void DropGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
};

namespace {

// Material constructors (materials.hxx): the record stores what the constructors leave behind.
DevMaterial ToDev(const pg3g_material &d) {
    DevMaterial m{};
    m.type = d.type;
    m.lightId = d.light_id;
    m.diffuse = mk(d.diffuse[0], d.diffuse[1], d.diffuse[2]);
    m.glossy = mk(d.glossy[0], d.glossy[1], d.glossy[2]);
    m.phongExp = d.phong_exponent;
    m.alpha = 0.1f;
    m.eta = 1.f;
    m.etaInv = 1.f;
    m.k = d.absorbance;
    const float inner = d.inner_ior, outer = d.outer_ior;
    switch (d.type) {
    case PG3G_MAT_LAMBERT:                                   // :253-266
        m.glossy = mk(0.f);
        m.sides = kSideFront;
        break;
    case PG3G_MAT_PHONG:                                     // :364-379
        m.diffuse = d.phong_diffuse_on ? m.diffuse : mk(0.f);
        m.glossy = d.phong_glossy_on ? m.glossy : mk(0.f);
        if (d.phong_glossy_on) m.diffuse = m.diffuse / 2;
        m.sides = kSideFront;
        break;
    case PG3G_MAT_GGX_CONDUCTOR:                             // :791-806
        m.eta = !IsTiny(outer) ? inner / outer : 1.0f;
        m.alpha = Clampf(d.alpha, 0.001f, 1.0f);
        m.sides = kSideFront;
        break;
    case PG3G_MAT_GGX_DIELECTRIC:                            // :996-1021
        if (!IsTiny(outer)) { m.eta = inner / outer; m.etaInv = outer / inner; }
        m.alpha = Clampf(d.alpha, 0.001f, 1.0f);
        m.sides = kSideFront | (d.back_side_light_sampling ? kSideBack : 0u);
        break;
    case PG3G_MAT_SMOOTH_CONDUCTOR:                          // :660-669
        m.eta = !IsTiny(outer) ? inner / outer : 1.0f;
        m.sides = 0;
        break;
    case PG3G_MAT_SMOOTH_DIELECTRIC:                         // :716-727
        m.eta = !IsTiny(outer) ? inner / outer : 1.0f;
        m.etaInv = !IsTiny(inner) ? outer / inner : 1.0f;
        m.sides = 0;
        break;
    }
    return m;
}

// Camera::Setup (camera.hxx:13-59), composed in FP64 and stored as FP32.
void CameraMatrix(const pg3g_camera_desc &c, float out[16]) {
    const double pi = 3.14159265358979323846;
    double f[3] = {c.forward[0], c.forward[1], c.forward[2]};
    const double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double &v : f) v /= fl;
    const double b[3] = {-f[0], -f[1], -f[2]};
    double up[3] = {c.up[1] * b[2] - c.up[2] * b[1], c.up[2] * b[0] - c.up[0] * b[2], c.up[0] * b[1] - c.up[1] * b[0]};
    const double ul = std::sqrt(up[0] * up[0] + up[1] * up[1] + up[2] * up[2]);
    for (double &v : up) v /= ul;
    const double left[3] = {b[1] * up[2] - b[2] * up[1], b[2] * up[0] - b[0] * up[2], b[0] * up[1] - b[1] * up[0]};
    const double P[3] = {c.position[0], c.position[1], c.position[2]};
    auto dot = [](const double *a, const double *q) { return a[0] * q[0] + a[1] * q[1] + a[2] * q[2]; };
    double M[4][4] = {{up[0], up[1], up[2], -dot(up, P)},
                      {left[0], left[1], left[2], -dot(left, P)},
                      {b[0], b[1], b[2], -dot(b, P)},
                      {0, 0, 0, 1}};
    const double half = 0.5 * c.hfov_deg * pi / 180.0;
    const double pixelSize = std::tan(half) / (0.5 * c.res_x);
    const double vhalf = std::atan(pixelSize * 0.5 * c.res_y);
    const double fx = 1.0 / std::tan(half), fy = 1.0 / std::tan(vhalf);
    const double n = 0.1, fr = 10000.0, d = 1.0 / (n - fr);
    const double Pm[4][4] = {{fx, 0, 0, 0}, {0, -fy, 0, 0}, {0, 0, (n + fr) * d, 2.0 * n * fr * d}, {0, 0, -1, 0}};
    double W[4][8];   // [P*M | I] reduced to [I | (P*M)^-1]
    for (int r = 0; r < 4; ++r)
        for (int col = 0; col < 4; ++col) {
            double s = 0;
            for (int i = 0; i < 4; ++i) s += Pm[r][i] * M[i][col];
            W[r][col] = s;
            W[r][col + 4] = (r == col) ? 1.0 : 0.0;
        }
    for (int col = 0; col < 4; ++col) {
        int piv = col;
        for (int r = col + 1; r < 4; ++r) if (std::fabs(W[r][col]) > std::fabs(W[piv][col])) piv = r;
        for (int k = 0; k < 8; ++k) std::swap(W[col][k], W[piv][k]);
        const double dv = W[col][col];
        for (int k = 0; k < 8; ++k) W[col][k] /= dv;
        for (int r = 0; r < 4; ++r) {
            if (r == col) continue;
            const double fct = W[r][col];
            for (int k = 0; k < 8; ++k) W[r][k] -= fct * W[col][k];
        }
    }
    // raster -> NDC: x' = 2x/res - 1, y' = 2y/res - 1, z' = 0
    const double S[4][4] = {{2.0 / c.res_x, 0, 0, -1}, {0, 2.0 / c.res_y, 0, -1}, {0, 0, 0, 0}, {0, 0, 0, 1}};
    for (int r = 0; r < 4; ++r)
        for (int col = 0; col < 4; ++col) {
            double s = 0;
            for (int i = 0; i < 4; ++i) s += W[r][i + 4] * S[i][col];
            out[r * 4 + col] = (float)s;
        }
}

template <class T>
hipError_t Upload(DevBuf &buf, const T *src, size_t count, hipStream_t s) {
    hipError_t e = buf.Alloc(sizeof(T) * std::max<size_t>(count, 1));
    if (e != hipSuccess || count == 0) return e;
    return hipMemcpyAsync(buf.p, src, sizeof(T) * count, hipMemcpyHostToDevice, s);
}

// Rectangles: triangles i and i + 1 of the scene (the two halves of a quad, as the reference's
// boxes list their walls) merge into one DevQuad when both lie in a plane x_A = const, their
// normals are bit-identical and exactly zero off axis A, their plane distances and materials are
// equal, they share a diagonal, and their four corners are the corners of an axis-aligned
// rectangle.  PG3G_TRIANGLES_ONLY=1 in the environment keeps every triangle (for comparisons).
bool MergeRectangles() {
    const char *e = std::getenv("PG3G_TRIANGLES_ONLY");
    return !(e && e[0] == '1');
}
bool RectangleOf(const pg3g_scene_desc *d, const std::vector<DevTri> &tris, uint32_t i, DevQuad &q, int &axis) {
    const DevTri &a = tris[i], &b = tris[i + 1];
    if (a.mat != b.mat || std::memcmp(a.n, b.n, sizeof(a.n)) != 0 || std::memcmp(&a.np0, &b.np0, 4) != 0) return false;
    axis = -1;
    for (int k = 0; k < 3; ++k) {
        if (a.n[k] == 0.f) continue;
        if (axis >= 0) return false;
        axis = k;
    }
    if (axis < 0) return false;
    const int u = (axis + 1) % 3, w = (axis + 2) % 3;
    float P[6][3];
    std::memcpy(P, d->triangles + (size_t)i * 9, sizeof(P));
    for (int j = 1; j < 6; ++j)
        if (P[j][axis] != P[0][axis]) return false;
    // corners: the distinct points of the two triangles
    float C[6][3];
    int nc = 0, shared = 0;
    for (int j = 0; j < 6; ++j) {
        bool seen = false;
        for (int c = 0; c < nc; ++c) seen |= std::memcmp(C[c], P[j], sizeof(C[c])) == 0;
        if (seen) { ++shared; continue; }
        std::memcpy(C[nc++], P[j], sizeof(C[0]));
    }
    if (nc != 4 || shared != 2) return false;
    const float lo0 = std::min(std::min(C[0][u], C[1][u]), std::min(C[2][u], C[3][u]));
    const float hi0 = std::max(std::max(C[0][u], C[1][u]), std::max(C[2][u], C[3][u]));
    const float lo1 = std::min(std::min(C[0][w], C[1][w]), std::min(C[2][w], C[3][w]));
    const float hi1 = std::max(std::max(C[0][w], C[1][w]), std::max(C[2][w], C[3][w]));
    if (!(lo0 < hi0 && lo1 < hi1)) return false;
    // each corner of [lo0, hi0] x [lo1, hi1] appears once
    int mask = 0;
    for (int c = 0; c < 4; ++c) {
        const bool uLo = C[c][u] == lo0, uHi = C[c][u] == hi0, wLo = C[c][w] == lo1, wHi = C[c][w] == hi1;
        if (!(uLo || uHi) || !(wLo || wHi)) return false;
        mask |= 1 << ((uHi ? 1 : 0) + (wHi ? 2 : 0));
    }
    if (mask != 15) return false;
    // the shared edge is a diagonal: its two points differ in both coordinates
    int sh[2], ns = 0;
    for (int j = 3; j < 6 && ns < 2; ++j)
        for (int k = 0; k < 3; ++k)
            if (std::memcmp(P[j], P[k], sizeof(P[0])) == 0) { sh[ns++] = j; break; }
    if (ns != 2 || P[sh[0]][u] == P[sh[1]][u] || P[sh[0]][w] == P[sh[1]][w]) return false;
    std::memset(&q, 0, sizeof(q));
    std::memcpy(q.n, a.n, sizeof(q.n));
    q.np0 = a.np0;
    q.lo0 = lo0; q.hi0 = hi0; q.lo1 = lo1; q.hi1 = hi1;
    q.mat = a.mat;
    return true;
}

int ValidateDesc(const pg3g_scene_desc *d) {
    if (!d) return Fail(PG3G_ERR_INVALID_ARG, "scene desc is null");
    if (d->num_triangles && (!d->triangles || !d->triangle_mat)) return Fail(PG3G_ERR_INVALID_ARG, "triangle arrays are null");
    if (d->num_spheres && (!d->spheres || !d->sphere_mat)) return Fail(PG3G_ERR_INVALID_ARG, "sphere arrays are null");
    if (d->num_materials == 0 || !d->materials) return Fail(PG3G_ERR_INVALID_ARG, "scene has no materials");
    if (d->num_lights && !d->lights) return Fail(PG3G_ERR_INVALID_ARG, "light array is null");
    for (uint32_t i = 0; i < d->num_triangles; ++i)
        if (d->triangle_mat[i] < 0 || (uint32_t)d->triangle_mat[i] >= d->num_materials)
            return Fail(PG3G_ERR_INVALID_ARG, "triangle material index out of range");
    for (uint32_t i = 0; i < d->num_spheres; ++i) {
        if (d->sphere_mat[i] < 0 || (uint32_t)d->sphere_mat[i] >= d->num_materials)
            return Fail(PG3G_ERR_INVALID_ARG, "sphere material index out of range");
        if (!(d->spheres[i * 4 + 3] > 0.f)) return Fail(PG3G_ERR_INVALID_ARG, "sphere radius must be positive");
    }
    for (uint32_t i = 0; i < d->num_materials; ++i) {
        const pg3g_material &m = d->materials[i];
        if (m.type < PG3G_MAT_LAMBERT || m.type > PG3G_MAT_SMOOTH_DIELECTRIC) return Fail(PG3G_ERR_INVALID_ARG, "unknown material type");
        if (m.light_id >= (int32_t)d->num_lights) return Fail(PG3G_ERR_INVALID_ARG, "material light id out of range");
        if (m.light_id >= 0 && d->lights[m.light_id].type != PG3G_LIGHT_AREA)
            return Fail(PG3G_ERR_INVALID_ARG, "only area lights can own geometry");
        const float vals[] = {m.diffuse[0], m.diffuse[1], m.diffuse[2], m.glossy[0], m.glossy[1], m.glossy[2],
                              m.alpha, m.inner_ior, m.outer_ior, m.absorbance, m.phong_exponent};
        for (float v : vals) if (!std::isfinite(v)) return Fail(PG3G_ERR_INVALID_ARG, "material parameter is not finite");
    }
    int bg = 0;
    for (uint32_t i = 0; i < d->num_lights; ++i) {
        const int t = d->lights[i].type;
        if (t < PG3G_LIGHT_AREA || t > PG3G_LIGHT_BACKGROUND) return Fail(PG3G_ERR_INVALID_ARG, "unknown light type");
        if (t == PG3G_LIGHT_BACKGROUND) ++bg;
    }
    if (bg > 1) return Fail(PG3G_ERR_INVALID_ARG, "at most one background light");
    if (d->num_lights > 1 && d->num_lights > 8)
        return Fail(PG3G_ERR_UNSUPPORTED, "more than 8 lights with light picking are not supported");
    if (d->em.width > 0) {
        if (!d->em.texels || d->em.height == 0) return Fail(PG3G_ERR_INVALID_ARG, "EM image arrays are null");
        if (d->em.sampler < PG3G_EM_COSINE || d->em.sampler > PG3G_EM_STEERABLE) return Fail(PG3G_ERR_INVALID_ARG, "unknown EM sampler");
    }
    if (d->camera.res_x == 0 || d->camera.res_y == 0) return Fail(PG3G_ERR_INVALID_ARG, "camera resolution is zero");
    if ((uint64_t)d->camera.res_x * d->camera.res_y > (1ull << 28)) return Fail(PG3G_ERR_INVALID_ARG, "resolution too large");
    return PG3G_OK;
}

int ValidateParams(const pg3g_render_params &p) {
    if (p.algorithm < PG3G_ALG_DIRECT_BSDF || p.algorithm > PG3G_ALG_PATH_TRACING)
        return Fail(PG3G_ERR_UNSUPPORTED, "algorithm is not on the GPU path (supported: dbs dlsa dlss dmis ptn pt)");
    if (p.algorithm == PG3G_ALG_PATH_TRACING) {
        if (!(p.splitting_budget >= 1.f && p.splitting_budget <= 64.f))
            return Fail(PG3G_ERR_UNSUPPORTED, "the GPU path tracer supports splitting budgets 1..64 (-sb)");
        if (!(p.splitting_level >= 0.f && p.splitting_level <= 1.f))
            return Fail(PG3G_ERR_INVALID_ARG, "splitting level must be in [0, 1]");
        if (!(p.light_to_bsdf_ratio > 0.f && p.light_to_bsdf_ratio <= 16.f))
            return Fail(PG3G_ERR_UNSUPPORTED, "the GPU path tracer supports light-to-BSDF sample ratios in (0, 16]");
        if (!(p.indirect_clipping >= 0.f)) return Fail(PG3G_ERR_INVALID_ARG, "indirect clipping must be >= 0");
    }
    if (p.max_path_length > 0 && p.min_path_length > p.max_path_length)
        return Fail(PG3G_ERR_INVALID_ARG, "min path length exceeds max path length");
    return PG3G_OK;
}

void ApplyParams(SceneView &v, const pg3g_render_params &p) {
    v.algorithm = p.algorithm;
    v.minPL = p.min_path_length;
    v.maxPL = p.max_path_length;
    v.iic = p.indirect_clipping;
    v.splitBudget = p.splitting_budget;
    v.splitLevel = p.splitting_level;
    v.lightRatio = p.light_to_bsdf_ratio;
    v.k0 = (uint32_t)p.seed;
    v.k1 = (uint32_t)(p.seed >> 32);
    v.sceneId = p.scene_id & 0xFFFFu;
}

SceneSlot *GetSlot(pg3g_context *ctx, uint32_t slot) {
    if (!ctx || slot >= ctx->slots.size() || !ctx->slots[slot] || !ctx->slots[slot]->valid) return nullptr;
    return ctx->slots[slot].get();
}

int UploadScene(pg3g_context *ctx, uint32_t slotIdx, const pg3g_scene_desc *d) {
    if (int rc = ValidateDesc(d)) return rc;
    if (int rc = ValidateParams(d->params)) return rc;
    PG3_HIP(hipSetDevice(ctx->device));
    if (slotIdx >= PG3G_MAX_SCENES) return Fail(PG3G_ERR_INVALID_ARG, "scene slot out of range");
    if (ctx->slots.size() <= slotIdx) ctx->slots.resize(slotIdx + 1);
    if (!ctx->slots[slotIdx]) ctx->slots[slotIdx].reset(new (std::nothrow) SceneSlot());
    if (!ctx->slots[slotIdx]) return Fail(PG3G_ERR_OUT_OF_MEMORY, "host allocation failed");
    SceneSlot &S = *ctx->slots[slotIdx];
    S.valid = false;
    hipStream_t st = ctx->stream;
    SceneView &v = S.view;
    v = SceneView{};

    std::vector<DevTri> tris(d->num_triangles);
    for (uint32_t i = 0; i < d->num_triangles; ++i) {
        DevTri &t = tris[i];
        std::memset(&t, 0, sizeof(t));
        const float *p = d->triangles + i * 9;
        const F3 P[3] = {mk(p[0], p[1], p[2]), mk(p[3], p[4], p[5]), mk(p[6], p[7], p[8])};
        const int ends[3][2] = {{2, 1}, {1, 0}, {0, 2}};   // edge k runs from P[i] to P[j]: (j, i)
        for (int e = 0; e < 3; ++e) {
            const F3 a = P[ends[e][0]], b = P[ends[e][1]];
            const F3 u = a - b, v = Cross(a, b);
            t.u[e][0] = u.x; t.u[e][1] = u.y; t.u[e][2] = u.z;
            t.v[e][0] = v.x; t.v[e][1] = v.y; t.v[e][2] = v.z;
        }
        const F3 n = Normalize(Cross(P[1] - P[0], P[2] - P[0]));
        t.n[0] = n.x; t.n[1] = n.y; t.n[2] = n.z;
        t.np0 = Dot(n, P[0]);
        t.mat = d->triangle_mat[i];
    }
    std::vector<DevQuad> quads[3];
    std::vector<DevTri> rest;
    if (MergeRectangles()) {
        for (uint32_t i = 0; i < d->num_triangles; ++i) {
            DevQuad q;
            int axis;
            if (i + 1 < d->num_triangles && RectangleOf(d, tris, i, q, axis)) {
                quads[axis].push_back(q);
                ++i;
            } else {
                rest.push_back(tris[i]);
            }
        }
    } else {
        rest = tris;
    }
    std::vector<DevQuad> allQuads;
    for (int a = 0; a < 3; ++a) {
        allQuads.insert(allQuads.end(), quads[a].begin(), quads[a].end());
        v.quadEnd[a] = (uint32_t)allQuads.size();
    }
    std::vector<DevSphere> sphs(d->num_spheres);
    for (uint32_t i = 0; i < d->num_spheres; ++i) {
        std::memset(&sphs[i], 0, sizeof(DevSphere));
        for (int k = 0; k < 4; ++k) (k < 3 ? sphs[i].c[k] : sphs[i].r) = d->spheres[i * 4 + k];
        sphs[i].mat = d->sphere_mat[i];
    }
    std::vector<DevMaterial> mats(d->num_materials);
    for (uint32_t i = 0; i < d->num_materials; ++i) mats[i] = ToDev(d->materials[i]);
    S.hostMats.assign(d->materials, d->materials + d->num_materials);
    S.evalHostMats = S.hostMats;
    S.diffMats.clear();
    S.numScalars = 0;
    S.numViews = 0;
    S.pendCount = 0;
    S.gradBaseHost.assign(d->num_materials, -1);
    std::vector<DevLight> lights(d->num_lights);
    v.bgLight = -1;
    for (uint32_t i = 0; i < d->num_lights; ++i) {
        const pg3g_light &l = d->lights[i];
        DevLight &L = lights[i];
        std::memset(&L, 0, sizeof(L));
        L.type = l.type;
        const F3 pw = mk(l.power[0], l.power[1], l.power[2]);
        if (l.type == PG3G_LIGHT_AREA) {                     // AreaLight ctor + SetPower, lights.hxx:62-112
            L.p0 = mk(l.p0[0], l.p0[1], l.p0[2]);
            L.e1 = mk(l.p1[0], l.p1[1], l.p1[2]) - L.p0;
            L.e2 = mk(l.p2[0], l.p2[1], l.p2[2]) - L.p0;
            const F3 n = Cross(L.e1, L.e2);
            const float len = Length(n);
            if (!(len > 0.f)) return Fail(PG3G_ERR_INVALID_ARG, "degenerate area light");
            L.invArea = 2.f / len;
            L.n = FrameFromZ(n / Length(n)).z;
            L.radiance = pw * (L.invArea / kPiF);
        } else if (l.type == PG3G_LIGHT_POINT) {             // PointLight::SetPower, lights.hxx:242-245
            L.p0 = mk(l.p0[0], l.p0[1], l.p0[2]);
            L.radiance = pw / (4 * kPiF);
        } else {
            L.radiance = pw;
            v.bgLight = (int32_t)i;
        }
    }
    PG3_HIP(Upload(S.quads, allQuads.data(), allQuads.size(), st));
    PG3_HIP(Upload(S.tris, rest.data(), rest.size(), st));
    PG3_HIP(Upload(S.sphs, sphs.data(), sphs.size(), st));
    PG3_HIP(Upload(S.mats, mats.data(), mats.size(), st));
    PG3_HIP(Upload(S.lights, lights.data(), lights.size(), st));
    v.quads = S.quads.as<DevQuad>();
    v.tris = S.tris.as<DevTri>();
    v.sphs = S.sphs.as<DevSphere>();
    v.mats = S.mats.as<DevMaterial>();
    v.lights = S.lights.as<DevLight>();
    v.numTris = (uint32_t)rest.size();
    v.numSphs = d->num_spheres;
    v.numLights = d->num_lights;
    v.matClasses = 1u << kClsMiss;
    for (const DevMaterial &m : mats) v.matClasses |= 1u << ClassOfType(m.type);
    PG3_HIP(Upload(S.gradBase, S.gradBaseHost.data(), S.gradBaseHost.size(), st));
    v.evalMats = v.mats;
    v.gradBase = S.gradBase.as<int32_t>();

    // Environment map + sampling tables
    DevEm &E = v.em;
    S.hasSteer = false;
    if (d->em.width > 0 && v.bgLight >= 0) {
        E.w = d->em.width;
        E.h = d->em.height;
        E.bilinear = d->em.bilinear ? 1 : 0;
        E.sampler = d->em.sampler;
        E.plan2ang = 1.0f / (2.0f * kPiF * kPiF);
        std::vector<float4> tex((size_t)E.w * E.h);
        for (size_t i = 0; i < tex.size(); ++i)
            tex[i] = make_float4(d->em.texels[i * 3], d->em.texels[i * 3 + 1], d->em.texels[i * 3 + 2], 0.f);
        PG3_HIP(Upload(S.texels, tex.data(), tex.size(), st));
        E.texels = S.texels.as<float4>();
        PG3_HIP(S.lum.Alloc(sizeof(float) * tex.size()));
        E.lum = S.lum.as<float>();
        PG3_HIP(LaunchEmLum(E.texels, S.lum.as<float>(), (uint32_t)tex.size(), st));
        S.hostTexels.assign(d->em.texels, d->em.texels + (size_t)E.w * E.h * 3);
        S.emDesc = d->em;
        S.emDesc.texels = nullptr;
        {
            // The 2-D CDF is built for every image EM: besides the simple sampler it serves
            // EnvironmentMap::EstimateIrradiance's MC fallback (em.hxx:123-170).
            PG3_HIP(S.margCdf.Alloc(sizeof(float) * (E.h + 1)));
            PG3_HIP(S.condCdf.Alloc(sizeof(float) * (size_t)E.h * (E.w + 1)));
            PG3_HIP(S.rowInt.Alloc(sizeof(float) * E.h));
            E.margCdf = S.margCdf.as<float>();
            E.condCdf = S.condCdf.as<float>();
            PG3_HIP(LaunchCdfBuild(E, S.condCdf.as<float>(), S.margCdf.as<float>(), S.rowInt.as<float>(), st));
        }
        if (E.sampler == PG3G_EM_STEERABLE) {
            std::string err;
            const uint32_t mn = d->em.st_min_subdiv_level >= 0 ? (uint32_t)d->em.st_min_subdiv_level : 4u;
            const uint32_t mx = d->em.st_max_subdiv_level >= 0 ? (uint32_t)d->em.st_max_subdiv_level : 9u;
            if (!BuildSteerableCached(d->em.texels, E.w, E.h, E.bilinear != 0, d->em.st_max_approx_error, mn, mx, S.steer,
                                      err))
                return Fail(PG3G_ERR_INVALID_ARG, err);
            S.hasSteer = true;
            PG3_HIP(Upload(S.nodeW, S.steer.nodeW.data(), S.steer.nodeW.size(), st));
            PG3_HIP(Upload(S.leafVerts, S.steer.leafVerts.data(), S.steer.leafVerts.size(), st));
            E.nodeW = S.nodeW.as<float>();
            E.leafVerts = S.leafVerts.as<float>();
            E.numLevels = S.steer.numLevels;
            for (uint32_t l = 0; l < S.steer.numLevels; ++l) {
                E.levelBase[l] = S.steer.levelBase[l];
                E.levelCount[l] = S.steer.levelCount[l];
            }
            for (int j = 0; j < 9; ++j) E.rootW[j] = S.steer.rootW[j];
        }
    }

    v.evalTexels = E.texels;
    CameraMatrix(d->camera, v.r2w);
    v.camPos = mk(d->camera.position[0], d->camera.position[1], d->camera.position[2]);
    v.resX = d->camera.res_x;
    v.resY = d->camera.res_y;
    S.params = d->params;
    ApplyParams(v, d->params);

    const size_t fbBytes = sizeof(double) * 3 * (size_t)v.resX * v.resY;
    PG3_HIP(S.fb.Alloc(fbBytes));
    PG3_HIP(hipMemsetAsync(S.fb.p, 0, fbBytes, st));
    S.iterations = 0;
    PG3_HIP(S.viewDev.Alloc(sizeof(SceneView)));
    PG3_HIP(hipMemcpyAsync(S.viewDev.p, &S.view, sizeof(SceneView), hipMemcpyHostToDevice, st));
    PG3_HIP(hipStreamSynchronize(st));
    S.valid = true;
    ++ctx->sceneEpoch;
    ++S.epoch;
    S.sampleValid = false;
    return PG3G_OK;
}

// ---- differentiable rendering ------------------------------------------------------------------
uint32_t DiffScalarCount(int32_t type) {
    return type == PG3G_MAT_LAMBERT ? 3u : (type == PG3G_MAT_GGX_CONDUCTOR ? 4u : 0u);
}

int SyncView(pg3g_context *ctx, SceneSlot &S) {
    PG3_HIP(hipMemcpyAsync(S.viewDev.p, &S.view, sizeof(SceneView), hipMemcpyHostToDevice, ctx->stream));
    PG3_HIP(hipStreamSynchronize(ctx->stream));
    ++ctx->sceneEpoch;
    ++S.epoch;
    return PG3G_OK;
}

// Rebuilds the EM sampling tables of a slot from S.hostTexels (already uploaded to S.texels).
int RebuildEmTables(pg3g_context *ctx, SceneSlot &S) {
    DevEm &E = S.view.em;
    PG3_HIP(LaunchEmLum(E.texels, S.lum.as<float>(), E.w * E.h, ctx->stream));
    PG3_HIP(LaunchCdfBuild(E, S.condCdf.as<float>(), S.margCdf.as<float>(), S.rowInt.as<float>(), ctx->stream));
    if (E.sampler == PG3G_EM_STEERABLE) {
        std::string err;
        const pg3g_em_desc &d = S.emDesc;
        const uint32_t mn = d.st_min_subdiv_level >= 0 ? (uint32_t)d.st_min_subdiv_level : 4u;
        const uint32_t mx = d.st_max_subdiv_level >= 0 ? (uint32_t)d.st_max_subdiv_level : 9u;
        SteerableTables t;
        if (!BuildSteerableCached(S.hostTexels.data(), E.w, E.h, E.bilinear != 0, d.st_max_approx_error, mn, mx, t, err))
            return Fail(PG3G_ERR_INVALID_ARG, err);
        S.steer = std::move(t);
        PG3_HIP(Upload(S.nodeW, S.steer.nodeW.data(), S.steer.nodeW.size(), ctx->stream));
        PG3_HIP(Upload(S.leafVerts, S.steer.leafVerts.data(), S.steer.leafVerts.size(), ctx->stream));
        E.nodeW = S.nodeW.as<float>();
        E.leafVerts = S.leafVerts.as<float>();
        E.numLevels = S.steer.numLevels;
        for (uint32_t l = 0; l < S.steer.numLevels; ++l) {
            E.levelBase[l] = S.steer.levelBase[l];
            E.levelCount[l] = S.steer.levelCount[l];
        }
        for (int j = 0; j < 9; ++j) E.rootW[j] = S.steer.rootW[j];
    }
    return PG3G_OK;
}

int CheckDiffSlot(pg3g_context *ctx, uint32_t slot, SceneSlot *&S) {
    S = GetSlot(ctx, slot);
    if (!S) return Fail(PG3G_ERR_NO_SCENE, "no scene in slot");
    if (S->view.algorithm != PG3G_ALG_PATH_TRACING || S->view.maxPL == 0 || S->view.splitBudget != 1.f ||
        S->view.iic > 0.f)
        return Fail(PG3G_ERR_UNSUPPORTED, "differentiable rendering needs the path tracer (pt) with -maxpl > 0 -sb 1 -iic 0");
    return PG3G_OK;
}

// ---- batches ---------------------------------------------------------------------------------
BatchLaunch MakeBatchLaunch(pg3g_batch *b) {
    BatchLaunch L;
    L.views = b->views.as<SceneView>();
    L.info = b->infoDev.as<BatchScene>();
    L.numScenes = (uint32_t)b->slots.size();
    L.numItems = b->numItems;
    L.work = b->work.as<uint32_t>();
    L.counters = b->ctx->counters.as<unsigned long long>();
    L.gridBlocks = b->gridBlocks;
    L.blockLanes = b->blockLanes;
    L.emKind = b->emKind;
    L.oneLight = b->oneLight;
    return L;
}

// Work items per scene for the current configuration (host table uploaded to the device).
int BuildBatchItems(pg3g_batch *b) {
    pg3g_context *ctx = b->ctx;
    uint64_t items = 0;
    for (size_t i = 0; i < b->slots.size(); ++i) {
        const SceneView &v = ctx->slots[b->slots[i]]->view;
        BatchScene &bs = b->info[i];
        bs.fb = ctx->slots[b->slots[i]]->fb.as<double>();
        bs.tilesX = (v.resX + 7u) / 8u;
        bs.numTiles = bs.tilesX * ((v.resY + 7u) / 8u);
        bs.spp = b->spp[i];
        bs.itersPerItem = std::min(b->cfg.iters_per_item, bs.spp);
        bs.itemBegin = (uint32_t)items;
        items += (uint64_t)bs.numTiles * ((bs.spp + bs.itersPerItem - 1u) / bs.itersPerItem);
        bs.pad = 0;
        if (items > 0xFFFFFFF0ull) return Fail(PG3G_ERR_INVALID_ARG, "batch has too many work items");
    }
    b->numItems = (uint32_t)items;
    // One persistent workgroup of up to 1024 lanes per compute unit (more workgroups when more
    // wavefronts per CU are asked for).
    const uint32_t waves = b->cfg.waves_per_cu;
    b->blockLanes = 64u * std::min(waves, 16u);
    b->gridBlocks = (uint32_t)std::max(1, ctx->numCUs) * std::max(1u, waves / 16u);
    // kBatchParts contiguous parts of the item sequence with about equal sample counts (the items
    // of one scene hold about equal sample counts)
    const uint32_t nS = (uint32_t)b->info.size();
    std::vector<double> w(nS);
    double total = 0.0;
    for (uint32_t i = 0; i < nS; ++i) {
        const uint32_t n = (i + 1 < nS ? b->info[i + 1].itemBegin : b->numItems) - b->info[i].itemBegin;
        w[i] = n ? 64.0 * b->info[i].numTiles * b->info[i].spp / n : 0.0;
        total += w[i] * n;
    }
    uint32_t parts[kBatchParts + 1];
    parts[0] = 0;
    for (uint32_t p = 1; p <= kBatchParts; ++p) parts[p] = b->numItems;
    {
        uint32_t p = 1;
        double acc = 0.0;
        for (uint32_t i = 0; i < nS && p < kBatchParts; ++i) {
            const uint32_t b0 = b->info[i].itemBegin, n = (i + 1 < nS ? b->info[i + 1].itemBegin : b->numItems) - b0;
            while (p < kBatchParts && w[i] > 0.0 && acc + w[i] * n >= total * p / kBatchParts) {
                const double need = total * p / kBatchParts - acc;
                parts[p] = std::max(parts[p - 1], b0 + std::min(n, (uint32_t)std::ceil(need / w[i])));
                ++p;
            }
            acc += w[i] * n;
        }
    }
    PG3_HIP(hipSetDevice(ctx->device));
    PG3_HIP(hipMemcpyAsync(b->work.as<uint32_t>() + kBatchParts + 1, parts, sizeof(parts), hipMemcpyHostToDevice,
                           ctx->stream));
    PG3_HIP(b->infoDev.Alloc(sizeof(BatchScene) * b->info.size()));
    PG3_HIP(hipMemcpyAsync(b->infoDev.p, b->info.data(), sizeof(BatchScene) * b->info.size(), hipMemcpyHostToDevice,
                           ctx->stream));
    PG3_HIP(hipStreamSynchronize(ctx->stream));
    b->DropGraph();
    return PG3G_OK;
}

int BatchCheck(pg3g_batch *b) {
    if (!b) return Fail(PG3G_ERR_INVALID_ARG, "batch is null");
    if (b->epoch != b->ctx->sceneEpoch)
        return Fail(PG3G_ERR_NO_SCENE, "the context's scenes changed after the batch was created");
    return PG3G_OK;
}

int BatchRenderPasses(pg3g_batch *b, uint32_t passes) {
    pg3g_context *ctx = b->ctx;
    PG3_HIP(hipSetDevice(ctx->device));
    const BatchLaunch L = MakeBatchLaunch(b);
    if (b->cfg.use_graph && !b->exec) {
        PG3_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const hipError_t le = LaunchBatchPass(L, ctx->stream);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
        PG3_HIP(le);
        PG3_HIP(ce);
        b->graph = g;
        PG3_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    for (uint32_t p = 0; p < passes; ++p) {
        if (b->cfg.use_graph) PG3_HIP(hipGraphLaunch(b->exec, ctx->stream));
        else PG3_HIP(LaunchBatchPass(L, ctx->stream));
    }
    for (size_t i = 0; i < b->slots.size(); ++i) ctx->slots[b->slots[i]]->iterations += (uint64_t)b->spp[i] * passes;
    b->passes += passes;
    return PG3G_OK;
}

int BatchTimed(pg3g_batch *b, uint32_t passes, float *ms) {
    pg3g_context *ctx = b->ctx;
    PG3_HIP(hipSetDevice(ctx->device));
    hipEvent_t e0, e1;
    PG3_HIP(hipEventCreate(&e0));
    PG3_HIP(hipEventCreate(&e1));
    hipError_t e = hipEventRecord(e0, ctx->stream);
    const int rc = (e == hipSuccess) ? BatchRenderPasses(b, passes) : PG3G_OK;
    if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != PG3G_OK) return rc;
    PG3_HIP(e);
    if (ms) *ms = t;
    return PG3G_OK;
}

// Clears the batch's framebuffers and restarts at pass 0.
int BatchRestart(pg3g_batch *b) {
    pg3g_context *ctx = b->ctx;
    PG3_HIP(hipSetDevice(ctx->device));
    for (size_t i = 0; i < b->slots.size(); ++i) {
        SceneSlot &S = *ctx->slots[b->slots[i]];
        PG3_HIP(hipMemsetAsync(S.fb.p, 0, S.fb.bytes, ctx->stream));
        S.iterations = 0;
    }
    PG3_HIP(hipMemsetAsync(b->work.p, 0, sizeof(uint32_t) * (kBatchParts + 1), ctx->stream));   // counters, pass
    PG3_HIP(hipStreamSynchronize(ctx->stream));
    b->passes = 0;
    return PG3G_OK;
}

// Host staging for the function-level entry points
struct Staging {
    pg3g_context *ctx;
    size_t next = 0;
    void *Get(size_t bytes, hipError_t &err) {
        DevBuf &b = ctx->scratch[next++];
        err = b.Alloc(bytes);
        return b.p;
    }
};

}  // namespace

extern "C" {

const char *pg3g_last_error(void) { return g_lastError.c_str(); }

int pg3g_create(pg3g_context **ctx, int device) {
    if (!ctx) return Fail(PG3G_ERR_INVALID_ARG, "ctx is null");
    *ctx = nullptr;
    int count = 0;
    PG3_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return Fail(PG3G_ERR_INVALID_ARG, "no such HIP device");
    PG3_HIP(hipSetDevice(device));
    std::unique_ptr<pg3g_context> c(new (std::nothrow) pg3g_context());
    if (!c) return Fail(PG3G_ERR_OUT_OF_MEMORY, "host allocation failed");
    c->device = device;
    PG3_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    PG3_HIP(hipDeviceGetAttribute(&c->numCUs, hipDeviceAttributeMultiprocessorCount, device));
    PG3_HIP(c->counters.Alloc(sizeof(unsigned long long) * 4));
    PG3_HIP(c->work.Alloc(sizeof(uint32_t)));
    // device address of the allocation-free adjoint's scratch array on this device (once per
    // device); a failure only disables pg3g_adjoint_noalloc
    c->noAllocReady = (InitNoAllocScratch(device) == hipSuccess);
    PG3_HIP(hipMemsetAsync(c->counters.p, 0, sizeof(unsigned long long) * 4, c->stream));
    PG3_HIP(hipStreamSynchronize(c->stream));
    *ctx = c.release();
    return PG3G_OK;
}

int pg3g_destroy(pg3g_context *ctx) {
    if (!ctx) return PG3G_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->slots.clear();
    DestroyWavefrontContext(ctx->wavefront);
    ctx->wavefront = nullptr;
    ctx->counters.Free();
    ctx->work.Free();
    for (auto &b : ctx->scratch) b.Free();
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PG3G_OK;
}

int pg3g_upload_scene(pg3g_context *ctx, const pg3g_scene_desc *desc) { return pg3g_upload_scene_slot(ctx, 0, desc); }

int pg3g_upload_scene_slot(pg3g_context *ctx, uint32_t slot, const pg3g_scene_desc *desc) {
    if (!ctx) return Fail(PG3G_ERR_INVALID_ARG, "ctx is null");
    return UploadScene(ctx, slot, desc);
}

}  // extern "C"

namespace {

// Launches iterations [first_iter, first_iter + n_iter) of a slot.
int RenderSlot(pg3g_context *ctx, SceneSlot *S, uint32_t first_iter, uint32_t n_iter) {
    if ((uint64_t)first_iter + n_iter > 0xFFFFFFFFull) return Fail(PG3G_ERR_INVALID_ARG, "iteration range overflows");
    PG3_HIP(hipSetDevice(ctx->device));
    if (ctx->kernelMode == PG3G_KERNEL_WAVEFRONT) {
        if (!ctx->wavefront) ctx->wavefront = CreateWavefrontContext();
        if (!ctx->wavefront) return Fail(PG3G_ERR_OUT_OF_MEMORY, "host allocation failed");
        PG3_HIP(LaunchWavefront(S->view, S->viewDev.as<SceneView>(), first_iter, n_iter, S->fb.as<double>(),
                                ctx->counters.as<unsigned long long>(), ctx->work.as<uint32_t>(), ctx->numCUs,
                                ctx->wavefront, ctx->stream));
    } else {
        PG3_HIP(LaunchMegakernel(S->view, S->viewDev.as<SceneView>(), first_iter, n_iter, S->fb.as<double>(),
                                 ctx->counters.as<unsigned long long>(), ctx->work.as<uint32_t>(), ctx->numCUs,
                                 ctx->stream));
    }
    S->iterations += n_iter;
    return PG3G_OK;
}

// Launches the iterations pg3g_run_iteration has queued for a slot.
int FlushQueued(pg3g_context *ctx, SceneSlot *S) {
    if (!S || S->pendCount == 0) return PG3G_OK;
    const uint32_t first = S->pendFirst, n = S->pendCount;
    S->pendCount = 0;
    return RenderSlot(ctx, S, first, n);
}

}  // namespace

extern "C" {

int pg3g_set_params(pg3g_context *ctx, uint32_t slot, const pg3g_render_params *p) {
    SceneSlot *S = GetSlot(ctx, slot);
    if (!S) return Fail(PG3G_ERR_NO_SCENE, "no scene in slot");
    if (!p) return Fail(PG3G_ERR_INVALID_ARG, "params is null");
    if (int rc = ValidateParams(*p)) return rc;
    if (int rc = FlushQueued(ctx, S)) return rc;               // queued iterations use the old parameters
    S->params = *p;
    ApplyParams(S->view, *p);
    PG3_HIP(hipSetDevice(ctx->device));
    return SyncView(ctx, *S);
}

int pg3g_set_kernel_mode(pg3g_context *ctx, int mode) {
    if (!ctx) return Fail(PG3G_ERR_INVALID_ARG, "ctx is null");
    if (mode != PG3G_KERNEL_MEGAKERNEL && mode != PG3G_KERNEL_WAVEFRONT) return Fail(PG3G_ERR_INVALID_ARG, "unknown kernel mode");
    ctx->kernelMode = mode;
    return PG3G_OK;
}

int pg3g_render(pg3g_context *ctx, uint32_t first_iter, uint32_t n_iter) { return pg3g_render_slot(ctx, 0, first_iter, n_iter); }

int pg3g_render_slot(pg3g_context *ctx, uint32_t slot, uint32_t first_iter, uint32_t n_iter) {
    SceneSlot *S = GetSlot(ctx, slot);
    if (!S) return Fail(PG3G_ERR_NO_SCENE, "no scene in slot");
    if (int rc = FlushQueued(ctx, S)) return rc;
    return RenderSlot(ctx, S, first_iter, n_iter);
}

int pg3g_run_iteration(pg3g_context *ctx, uint32_t slot, uint32_t iteration, int launch_now) {
    SceneSlot *S = GetSlot(ctx, slot);
    if (!S) return Fail(PG3G_ERR_NO_SCENE, "no scene in slot");
    if (S->pendCount > 0 && (uint64_t)iteration != (uint64_t)S->pendFirst + S->pendCount)
        if (int rc = FlushQueued(ctx, S)) return rc;            // not the next iteration: launch the run
    if (S->pendCount == 0) S->pendFirst = iteration;
    ++S->pendCount;
    if (launch_now || S->pendCount >= PG3G_ITERATION_RUN) return FlushQueued(ctx, S);
    return PG3G_OK;
}

int pg3g_render_timed(pg3g_context *ctx, uint32_t slot, uint32_t first_iter, uint32_t n_iter, float *elapsed_ms) {
    if (!ctx) return Fail(PG3G_ERR_INVALID_ARG, "ctx is null");
    PG3_HIP(hipSetDevice(ctx->device));
    hipEvent_t e0, e1;
    PG3_HIP(hipEventCreate(&e0));
    PG3_HIP(hipEventCreate(&e1));
    PG3_HIP(hipEventRecord(e0, ctx->stream));
    const int rc = pg3g_render_slot(ctx, slot, first_iter, n_iter);
    hipError_t e = hipEventRecord(e1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != PG3G_OK) return rc;
    PG3_HIP(e);
    if (elapsed_ms) *elapsed_ms = ms;
    return PG3G_OK;
}

int pg3g_read_framebuffer(pg3g_context *ctx, double *sum_rgb, uint64_t *iterations) {
    return pg3g_read_framebuffer_slot(ctx, 0, sum_rgb, iterations);
}

int pg3g_read_framebuffer_slot(pg3g_context *ctx, uint32_t slot, double *sum_rgb, uint64_t *iterations) {
    SceneSlot *S = GetSlot(ctx, slot);
    if (!S) return Fail(PG3G_ERR_NO_SCENE, "no scene in slot");
    if (int rc = FlushQueued(ctx, S)) return rc;
    PG3_HIP(hipSetDevice(ctx->device));
    if (sum_rgb) PG3_HIP(hipMemcpyAsync(sum_rgb, S->fb.p, sizeof(double) * 3 * (size_t)S->view.resX * S->view.resY,
                                        hipMemcpyDeviceToHost, ctx->stream));
    PG3_HIP(hipStreamSynchronize(ctx->stream));
    if (iterations) *iterations = S->iterations;
    return PG3G_OK;
}

int pg3g_clear_framebuffer(pg3g_context *ctx, uint32_t slot) {
    SceneSlot *S = GetSlot(ctx, slot);
    if (!S) return Fail(PG3G_ERR_NO_SCENE, "no scene in slot");
    S->pendCount = 0;                                             // queued iterations are dropped
    PG3_HIP(hipSetDevice(ctx->device));
    PG3_HIP(hipMemsetAsync(S->fb.p, 0, S->fb.bytes, ctx->stream));
    S->iterations = 0;
    return PG3G_OK;
}

int pg3g_framebuffer_device_ptr(pg3g_context *ctx, uint32_t slot, double **fb_dev) {
    SceneSlot *S = GetSlot(ctx, slot);
    if (!S) return Fail(PG3G_ERR_NO_SCENE, "no scene in slot");
    if (!fb_dev) return Fail(PG3G_ERR_INVALID_ARG, "fb_dev is null");
    if (int rc = FlushQueued(ctx, S)) return rc;
    *fb_dev = S->fb.as<double>();
    return PG3G_OK;
}

int pg3g_synchronize(pg3g_context *ctx) {
    if (!ctx) return Fail(PG3G_ERR_INVALID_ARG, "ctx is null");
    for (auto &slot : ctx->slots)
        if (slot && slot->valid)
            if (int rc = FlushQueued(ctx, slot.get())) return rc;
    PG3_HIP(hipSetDevice(ctx->device));
    PG3_HIP(hipStreamSynchronize(ctx->stream));
    return PG3G_OK;
}

int pg3g_read_counters(pg3g_context *ctx, uint64_t *segments, uint64_t *camera_samples) {
    if (!ctx) return Fail(PG3G_ERR_INVALID_ARG, "ctx is null");
    unsigned long long h[2] = {0, 0};
    PG3_HIP(hipSetDevice(ctx->device));
    PG3_HIP(hipMemcpyAsync(h, ctx->counters.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    PG3_HIP(hipStreamSynchronize(ctx->stream));
    if (segments) *segments = h[0];
    if (camera_samples) *camera_samples = h[1];
    return PG3G_OK;
}

int pg3g_read_ray_counters(pg3g_context *ctx, uint64_t *extension_rays, uint64_t *shadow_rays) {
    if (!ctx) return Fail(PG3G_ERR_INVALID_ARG, "ctx is null");
    unsigned long long h[4] = {0, 0, 0, 0};
    PG3_HIP(hipSetDevice(ctx->device));
    PG3_HIP(hipMemcpyAsync(h, ctx->counters.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    PG3_HIP(hipStreamSynchronize(ctx->stream));
    if (extension_rays) *extension_rays = h[2];
    if (shadow_rays) *shadow_rays = h[3];
    return PG3G_OK;
}

int pg3g_scene_primitives(pg3g_context *ctx, uint32_t slot, uint32_t *rectangles, uint32_t *triangles,
                          uint32_t *spheres) {
    if (!ctx) return Fail(PG3G_ERR_INVALID_ARG, "ctx is null");
    SceneSlot *S = GetSlot(ctx, slot);
    if (!S) return Fail(PG3G_ERR_NO_SCENE, "no scene uploaded in this slot");
    if (rectangles) *rectangles = S->view.quadEnd[2];
    if (triangles) *triangles = S->view.numTris;
    if (spheres) *spheres = S->view.numSphs;
    return PG3G_OK;
}

int pg3g_reset_counters(pg3g_context *ctx) {
    if (!ctx) return Fail(PG3G_ERR_INVALID_ARG, "ctx is null");
    PG3_HIP(hipSetDevice(ctx->device));
    PG3_HIP(hipMemsetAsync(ctx->counters.p, 0, ctx->counters.bytes, ctx->stream));
    return PG3G_OK;
}

// ---- function-level entry points ----------------------------------------------------------------
int pg3g_bsdf_eval_batch(pg3g_context *ctx, const pg3g_material *mat, uint32_t n, const float *wil, const float *wol,
                         float *f, float *pdf, float *comp_prob) {
    if (!ctx || !mat || (n && (!wil || !wol || !f || !pdf || !comp_prob))) return Fail(PG3G_ERR_INVALID_ARG, "null argument");
    if (n == 0) return PG3G_OK;
    PG3_HIP(hipSetDevice(ctx->device));
    Staging sg{ctx};
    hipError_t e1, e2, e3, e4, e5;
    float *dWil = (float *)sg.Get(12ull * n, e1), *dWol = (float *)sg.Get(12ull * n, e2), *dF = (float *)sg.Get(12ull * n, e3);
    float *dPdf = (float *)sg.Get(4ull * n, e4), *dComp = (float *)sg.Get(4ull * n, e5);
    PG3_HIP(e1); PG3_HIP(e2); PG3_HIP(e3); PG3_HIP(e4); PG3_HIP(e5);
    PG3_HIP(hipMemcpyAsync(dWil, wil, 12ull * n, hipMemcpyHostToDevice, ctx->stream));
    PG3_HIP(hipMemcpyAsync(dWol, wol, 12ull * n, hipMemcpyHostToDevice, ctx->stream));
    PG3_HIP(LaunchBsdfEval(ToDev(*mat), n, dWil, dWol, dF, dPdf, dComp, ctx->stream));
    PG3_HIP(hipMemcpyAsync(f, dF, 12ull * n, hipMemcpyDeviceToHost, ctx->stream));
    PG3_HIP(hipMemcpyAsync(pdf, dPdf, 4ull * n, hipMemcpyDeviceToHost, ctx->stream));
    PG3_HIP(hipMemcpyAsync(comp_prob, dComp, 4ull * n, hipMemcpyDeviceToHost, ctx->stream));
    PG3_HIP(hipStreamSynchronize(ctx->stream));
    return PG3G_OK;
}

int pg3g_bsdf_sample_batch(pg3g_context *ctx, const pg3g_material *mat, uint32_t n, const float *wol, const float *u,
                           float *wil, float *f, float *pdf, float *comp_prob) {
    if (!ctx || !mat || (n && (!wol || !u || !wil || !f || !pdf || !comp_prob))) return Fail(PG3G_ERR_INVALID_ARG, "null argument");
    if (n == 0) return PG3G_OK;
    PG3_HIP(hipSetDevice(ctx->device));
    Staging sg{ctx};
    hipError_t e[6];
    float *dWol = (float *)sg.Get(12ull * n, e[0]), *dU = (float *)sg.Get(12ull * n, e[1]), *dWil = (float *)sg.Get(12ull * n, e[2]);
    float *dF = (float *)sg.Get(12ull * n, e[3]), *dPdf = (float *)sg.Get(4ull * n, e[4]), *dComp = (float *)sg.Get(4ull * n, e[5]);
    for (hipError_t x : e) PG3_HIP(x);
    PG3_HIP(hipMemcpyAsync(dWol, wol, 12ull * n, hipMemcpyHostToDevice, ctx->stream));
    PG3_HIP(hipMemcpyAsync(dU, u, 12ull * n, hipMemcpyHostToDevice, ctx->stream));
    PG3_HIP(LaunchBsdfSample(ToDev(*mat), n, dWol, dU, dWil, dF, dPdf, dComp, ctx->stream));
    PG3_HIP(hipMemcpyAsync(wil, dWil, 12ull * n, hipMemcpyDeviceToHost, ctx->stream));
    PG3_HIP(hipMemcpyAsync(f, dF, 12ull * n, hipMemcpyDeviceToHost, ctx->stream));
    PG3_HIP(hipMemcpyAsync(pdf, dPdf, 4ull * n, hipMemcpyDeviceToHost, ctx->stream));
    PG3_HIP(hipMemcpyAsync(comp_prob, dComp, 4ull * n, hipMemcpyDeviceToHost, ctx->stream));
    PG3_HIP(hipStreamSynchronize(ctx->stream));
    return PG3G_OK;
}

int pg3g_em_sample_batch(pg3g_context *ctx, uint32_t n, const float *normals, const int32_t *side_flags, const float *u,
                         float *dir, float *pdf, float *radiance_cos, int32_t *ok) {
    SceneSlot *S = GetSlot(ctx, 0);
    if (!S) return Fail(PG3G_ERR_NO_SCENE, "no scene in slot 0");
    if (S->view.em.w == 0) return Fail(PG3G_ERR_INVALID_ARG, "scene has no EM image");
    if (n && (!normals || !side_flags || !u || !dir || !pdf || !radiance_cos || !ok)) return Fail(PG3G_ERR_INVALID_ARG, "null argument");
    if (n == 0) return PG3G_OK;
    PG3_HIP(hipSetDevice(ctx->device));
    Staging sg{ctx};
    hipError_t e[7];
    float *dN = (float *)sg.Get(12ull * n, e[0]);
    int32_t *dS = (int32_t *)sg.Get(4ull * n, e[1]);
    float *dU = (float *)sg.Get(12ull * n, e[2]), *dD = (float *)sg.Get(12ull * n, e[3]), *dP = (float *)sg.Get(4ull * n, e[4]);
    float *dR = (float *)sg.Get(12ull * n, e[5]);
    int32_t *dOk = (int32_t *)sg.Get(4ull * n, e[6]);
    for (hipError_t x : e) PG3_HIP(x);
    PG3_HIP(hipMemcpyAsync(dN, normals, 12ull * n, hipMemcpyHostToDevice, ctx->stream));
    PG3_HIP(hipMemcpyAsync(dS, side_flags, 4ull * n, hipMemcpyHostToDevice, ctx->stream));
    PG3_HIP(hipMemcpyAsync(dU, u, 12ull * n, hipMemcpyHostToDevice, ctx->stream));
    PG3_HIP(LaunchEmSample(S->view, n, dN, dS, dU, dD, dP, dR, dOk, ctx->stream));
    PG3_HIP(hipMemcpyAsync(dir, dD, 12ull * n, hipMemcpyDeviceToHost, ctx->stream));
    PG3_HIP(hipMemcpyAsync(pdf, dP, 4ull * n, hipMemcpyDeviceToHost, ctx->stream));
    PG3_HIP(hipMemcpyAsync(radiance_cos, dR, 12ull * n, hipMemcpyDeviceToHost, ctx->stream));
    PG3_HIP(hipMemcpyAsync(ok, dOk, 4ull * n, hipMemcpyDeviceToHost, ctx->stream));
    PG3_HIP(hipStreamSynchronize(ctx->stream));
    return PG3G_OK;
}

int pg3g_em_pdf_batch(pg3g_context *ctx, uint32_t n, const float *dirs, const float *normals, const int32_t *side_flags,
                      float *pdf) {
    SceneSlot *S = GetSlot(ctx, 0);
    if (!S) return Fail(PG3G_ERR_NO_SCENE, "no scene in slot 0");
    if (S->view.em.w == 0) return Fail(PG3G_ERR_INVALID_ARG, "scene has no EM image");
    if (n && (!dirs || !normals || !side_flags || !pdf)) return Fail(PG3G_ERR_INVALID_ARG, "null argument");
    if (n == 0) return PG3G_OK;
    PG3_HIP(hipSetDevice(ctx->device));
    Staging sg{ctx};
    hipError_t e[4];
    float *dD = (float *)sg.Get(12ull * n, e[0]), *dN = (float *)sg.Get(12ull * n, e[1]);
    int32_t *dS = (int32_t *)sg.Get(4ull * n, e[2]);
    float *dP = (float *)sg.Get(4ull * n, e[3]);
    for (hipError_t x : e) PG3_HIP(x);
    PG3_HIP(hipMemcpyAsync(dD, dirs, 12ull * n, hipMemcpyHostToDevice, ctx->stream));
    PG3_HIP(hipMemcpyAsync(dN, normals, 12ull * n, hipMemcpyHostToDevice, ctx->stream));
    PG3_HIP(hipMemcpyAsync(dS, side_flags, 4ull * n, hipMemcpyHostToDevice, ctx->stream));
    PG3_HIP(LaunchEmPdf(S->view, n, dD, dN, dS, dP, ctx->stream));
    PG3_HIP(hipMemcpyAsync(pdf, dP, 4ull * n, hipMemcpyDeviceToHost, ctx->stream));
    PG3_HIP(hipStreamSynchronize(ctx->stream));
    return PG3G_OK;
}

int pg3g_em_eval_batch(pg3g_context *ctx, uint32_t n, const float *dirs, float *rgb) {
    SceneSlot *S = GetSlot(ctx, 0);
    if (!S) return Fail(PG3G_ERR_NO_SCENE, "no scene in slot 0");
    if (S->view.em.w == 0) return Fail(PG3G_ERR_INVALID_ARG, "scene has no EM image");
    if (n && (!dirs || !rgb)) return Fail(PG3G_ERR_INVALID_ARG, "null argument");
    if (n == 0) return PG3G_OK;
    PG3_HIP(hipSetDevice(ctx->device));
    Staging sg{ctx};
    hipError_t e[2];
    float *dD = (float *)sg.Get(12ull * n, e[0]), *dR = (float *)sg.Get(12ull * n, e[1]);
    for (hipError_t x : e) PG3_HIP(x);
    PG3_HIP(hipMemcpyAsync(dD, dirs, 12ull * n, hipMemcpyHostToDevice, ctx->stream));
    PG3_HIP(LaunchEmEval(S->view, n, dD, dR, ctx->stream));
    PG3_HIP(hipMemcpyAsync(rgb, dR, 12ull * n, hipMemcpyDeviceToHost, ctx->stream));
    PG3_HIP(hipStreamSynchronize(ctx->stream));
    return PG3G_OK;
}

int pg3g_intersect_batch(pg3g_context *ctx, uint32_t n, const float *org, const float *dir, const float *tmin,
                         float *dist, float *normal, int32_t *mat) {
    SceneSlot *S = GetSlot(ctx, 0);
    if (!S) return Fail(PG3G_ERR_NO_SCENE, "no scene in slot 0");
    if (n && (!org || !dir || !tmin || !dist || !normal || !mat)) return Fail(PG3G_ERR_INVALID_ARG, "null argument");
    if (n == 0) return PG3G_OK;
    PG3_HIP(hipSetDevice(ctx->device));
    Staging sg{ctx};
    hipError_t e[6];
    float *dO = (float *)sg.Get(12ull * n, e[0]), *dD = (float *)sg.Get(12ull * n, e[1]);
    float *dT = (float *)sg.Get(4ull * n, e[2]), *dR = (float *)sg.Get(4ull * n, e[3]);
    float *dN = (float *)sg.Get(12ull * n, e[4]);
    int32_t *dM = (int32_t *)sg.Get(4ull * n, e[5]);
    for (hipError_t x : e) PG3_HIP(x);
    PG3_HIP(hipMemcpyAsync(dO, org, 12ull * n, hipMemcpyHostToDevice, ctx->stream));
    PG3_HIP(hipMemcpyAsync(dD, dir, 12ull * n, hipMemcpyHostToDevice, ctx->stream));
    PG3_HIP(hipMemcpyAsync(dT, tmin, 4ull * n, hipMemcpyHostToDevice, ctx->stream));
    PG3_HIP(LaunchIntersect(S->view, n, dO, dD, dT, dR, dN, dM, ctx->stream));
    PG3_HIP(hipMemcpyAsync(dist, dR, 4ull * n, hipMemcpyDeviceToHost, ctx->stream));
    PG3_HIP(hipMemcpyAsync(normal, dN, 12ull * n, hipMemcpyDeviceToHost, ctx->stream));
    PG3_HIP(hipMemcpyAsync(mat, dM, 4ull * n, hipMemcpyDeviceToHost, ctx->stream));
    PG3_HIP(hipStreamSynchronize(ctx->stream));
    return PG3G_OK;
}

int pg3g_occluded_batch(pg3g_context *ctx, uint32_t n, const float *p, const float *dir, const float *tmax,
                        int32_t *occluded) {
    SceneSlot *S = GetSlot(ctx, 0);
    if (!S) return Fail(PG3G_ERR_NO_SCENE, "no scene in slot 0");
    if (n && (!p || !dir || !tmax || !occluded)) return Fail(PG3G_ERR_INVALID_ARG, "null argument");
    if (n == 0) return PG3G_OK;
    PG3_HIP(hipSetDevice(ctx->device));
    Staging sg{ctx};
    hipError_t e[4];
    float *dP = (float *)sg.Get(12ull * n, e[0]), *dD = (float *)sg.Get(12ull * n, e[1]);
    float *dT = (float *)sg.Get(4ull * n, e[2]);
    int32_t *dO = (int32_t *)sg.Get(4ull * n, e[3]);
    for (hipError_t x : e) PG3_HIP(x);
    PG3_HIP(hipMemcpyAsync(dP, p, 12ull * n, hipMemcpyHostToDevice, ctx->stream));
    PG3_HIP(hipMemcpyAsync(dD, dir, 12ull * n, hipMemcpyHostToDevice, ctx->stream));
    PG3_HIP(hipMemcpyAsync(dT, tmax, 4ull * n, hipMemcpyHostToDevice, ctx->stream));
    PG3_HIP(LaunchOccluded(S->view, n, dP, dD, dT, dO, ctx->stream));
    PG3_HIP(hipMemcpyAsync(occluded, dO, 4ull * n, hipMemcpyDeviceToHost, ctx->stream));
    PG3_HIP(hipStreamSynchronize(ctx->stream));
    return PG3G_OK;
}

int pg3g_steerable_info(pg3g_context *ctx, uint32_t *num_leaves, uint32_t *num_vertices, uint32_t *num_levels) {
    SceneSlot *S = GetSlot(ctx, 0);
    if (!S || !S->hasSteer) return Fail(PG3G_ERR_NO_SCENE, "slot 0 has no steerable sampler");
    if (num_leaves) *num_leaves = S->steer.numLeaves;
    if (num_vertices) *num_vertices = S->steer.numVertices;
    if (num_levels) *num_levels = S->steer.numLevels;
    return PG3G_OK;
}

int pg3g_steerable_leaves(pg3g_context *ctx, float *leaf_dirs) {
    SceneSlot *S = GetSlot(ctx, 0);
    if (!S || !S->hasSteer) return Fail(PG3G_ERR_NO_SCENE, "slot 0 has no steerable sampler");
    if (!leaf_dirs) return Fail(PG3G_ERR_INVALID_ARG, "null argument");
    for (uint32_t i = 0; i < S->steer.numLeaves; ++i)
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) leaf_dirs[i * 9 + k * 3 + c] = S->steer.leafVerts[(size_t)i * 36 + k * 12 + c];
    return PG3G_OK;
}

int pg3g_steerable_root_weight(pg3g_context *ctx, float *weight9) {
    SceneSlot *S = GetSlot(ctx, 0);
    if (!S || !S->hasSteer) return Fail(PG3G_ERR_NO_SCENE, "slot 0 has no steerable sampler");
    if (!weight9) return Fail(PG3G_ERR_INVALID_ARG, "null argument");
    for (int j = 0; j < 9; ++j) weight9[j] = S->steer.rootW[j];
    return PG3G_OK;
}

int pg3g_steerable_build(const pg3g_em_desc *em, uint32_t num_threads, pg3g_steerable_build_info *info,
                         float *leaf_dirs, uint32_t leaf_capacity, float *root_weight9) {
    if (!em || !em->texels || em->width == 0 || em->height == 0) return Fail(PG3G_ERR_INVALID_ARG, "no EM image");
    const uint32_t mn = em->st_min_subdiv_level >= 0 ? (uint32_t)em->st_min_subdiv_level : 4u;
    const uint32_t mx = em->st_max_subdiv_level >= 0 ? (uint32_t)em->st_max_subdiv_level : 9u;
    const uint32_t threads = num_threads ? num_threads : SteerableBuildThreads();
    SteerableTables t;
    std::string err;
    const auto t0 = std::chrono::steady_clock::now();
    bool cached = false;
    if (!BuildSteerableCached(em->texels, em->width, em->height, em->bilinear != 0, em->st_max_approx_error, mn, mx, t,
                              err, threads, &cached))
        return Fail(PG3G_ERR_INVALID_ARG, err);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (info) {
        info->num_leaves = t.numLeaves;
        info->num_vertices = t.numVertices;
        info->num_levels = t.numLevels;
        info->threads = threads;
        info->from_cache = cached ? 1u : 0u;
        info->build_ms = ms;
    }
    if (leaf_dirs) {
        if (leaf_capacity < t.numLeaves) return Fail(PG3G_ERR_INVALID_ARG, "leaf_dirs holds fewer leaves than the build");
        for (uint32_t i = 0; i < t.numLeaves; ++i)
            for (int k = 0; k < 3; ++k)
                for (int c = 0; c < 3; ++c) leaf_dirs[(size_t)i * 9 + k * 3 + c] = t.leafVerts[(size_t)i * 36 + k * 12 + c];
    }
    if (root_weight9)
        for (int j = 0; j < 9; ++j) root_weight9[j] = t.rootW[j];
    return PG3G_OK;
}

int pg3g_steerable_cache_dir(const char *dir) {
    SetSteerableCacheDir(dir);
    return PG3G_OK;
}

int pg3g_load_image(const char *path, uint32_t *width, uint32_t *height, float *rgb, uint64_t capacity) {
    if (!path || !width || !height) return Fail(PG3G_ERR_INVALID_ARG, "null argument");
    std::vector<float> img;
    uint32_t w = 0, h = 0;
    std::string err;
    if (!ReadImageFile(path, img, w, h, err)) return Fail(PG3G_ERR_INVALID_ARG, err);
    *width = w;
    *height = h;
    if (rgb) {
        if (capacity < img.size()) return Fail(PG3G_ERR_INVALID_ARG, "image buffer too small");
        std::memcpy(rgb, img.data(), img.size() * sizeof(float));
    }
    return PG3G_OK;
}

int pg3g_load_em(const char *path, float azimuth_rotation, float scale, uint32_t *width, uint32_t *height, float *rgb,
                 uint64_t capacity) {
    if (!path || !width || !height) return Fail(PG3G_ERR_INVALID_ARG, "null argument");
    std::vector<float> img;
    uint32_t w = 0, h = 0;
    std::string err;
    if (!LoadEmFile(path, azimuth_rotation, scale, img, w, h, err)) return Fail(PG3G_ERR_INVALID_ARG, err);
    *width = w;
    *height = h;
    if (rgb) {
        if (capacity < img.size()) return Fail(PG3G_ERR_INVALID_ARG, "image buffer too small");
        std::memcpy(rgb, img.data(), img.size() * sizeof(float));
    }
    return PG3G_OK;
}

int pg3g_save_image(const char *path, const double *rgb, uint32_t width, uint32_t height, double scale, float gamma) {
    if (!path || !rgb || width == 0 || height == 0) return Fail(PG3G_ERR_INVALID_ARG, "bad image arguments");
    std::string err;
    if (!WriteImageFile(path, rgb, width, height, scale, gamma, err)) return Fail(PG3G_ERR_INVALID_ARG, err);
    return PG3G_OK;
}

int pg3g_image_compare(const float *image, const float *reference, uint32_t width, uint32_t height,
                       pg3g_image_diff *out, float *diff_rgb) {
    if (!image || !reference || !out || width == 0 || height == 0) return Fail(PG3G_ERR_INVALID_ARG, "bad arguments");
    const size_t n = (size_t)width * height;
    auto lum = [](const float *p) { return 0.212671 * p[0] + 0.715160 * p[1] + 0.072169 * p[2]; };
    double meanRef = 0, meanImg = 0;
    for (size_t i = 0; i < n; ++i) {
        meanRef += std::fabs(lum(reference + i * 3));
        meanImg += std::fabs(lum(image + i * 3));
    }
    meanRef /= (double)n;
    meanImg /= (double)n;
    const double floor = std::max(1e-3 * meanRef, 1e-12);
    double sumRel2 = 0, sumAbs2 = 0, sumRef2 = 0, maxRel = 0;
    for (size_t i = 0; i < n; ++i) {
        const double a = lum(image + i * 3), b = lum(reference + i * 3), d = a - b;
        const double rel = d / std::max(std::fabs(b), floor);
        sumRel2 += rel * rel;
        sumAbs2 += d * d;
        sumRef2 += b * b;
        maxRel = std::max(maxRel, std::fabs(rel));
        if (diff_rgb)
            for (int c = 0; c < 3; ++c) diff_rgb[i * 3 + c] = std::fabs(image[i * 3 + c] - reference[i * 3 + c]);
    }
    out->pixel_rel_rmse = std::sqrt(sumRel2 / (double)n);
    out->image_rel_rmse = sumRef2 > 0 ? std::sqrt(sumAbs2 / sumRef2) : std::sqrt(sumAbs2 / (double)n);
    out->max_pixel_rel = maxRel;
    out->mean_luminance = meanImg;
    out->mean_luminance_reference = meanRef;
    return PG3G_OK;
}

int pg3g_batch_create(pg3g_context *ctx, uint32_t num_scenes, const uint32_t *slots, const uint32_t *spp,
                      pg3g_batch **batch) {
    if (!ctx || !slots || !spp || !batch) return Fail(PG3G_ERR_INVALID_ARG, "null argument");
    *batch = nullptr;
    if (num_scenes == 0 || num_scenes > PG3G_MAX_SCENES) return Fail(PG3G_ERR_INVALID_ARG, "bad scene count");
    std::unique_ptr<pg3g_batch> b(new (std::nothrow) pg3g_batch());
    if (!b) return Fail(PG3G_ERR_OUT_OF_MEMORY, "host allocation failed");
    b->ctx = ctx;
    b->epoch = ctx->sceneEpoch;
    std::vector<SceneView> views;
    int em = -100, one = -100;
    for (uint32_t i = 0; i < num_scenes; ++i) {
        SceneSlot *S = GetSlot(ctx, slots[i]);
        if (!S) return Fail(PG3G_ERR_NO_SCENE, "batch names an empty slot");
        if (spp[i] == 0) return Fail(PG3G_ERR_INVALID_ARG, "spp must be >= 1");
        if (S->view.algorithm != PG3G_ALG_PATH_TRACING || S->view.splitBudget != 1.f || S->view.iic > 0.f)
            return Fail(PG3G_ERR_UNSUPPORTED, "batched rendering supports the path tracer (pt) with -sb 1 -iic 0");
        for (uint32_t j = 0; j < i; ++j)
            if (slots[j] == slots[i]) return Fail(PG3G_ERR_INVALID_ARG, "a slot appears twice in the batch");
        const LightSpec ls = SceneLightSpec(S->view);
        const int o = (ls.numLights == 1) ? 1 : 0;
        em = (em == -100 || em == ls.em) ? ls.em : -200;
        one = (one == -100 || one == o) ? o : -200;
        views.push_back(S->view);
        b->slots.push_back(slots[i]);
        b->spp.push_back(spp[i]);
    }
    b->emKind = (em == -200 || one == -200) ? 0 : em;
    b->oneLight = (em == -200 || one == -200) ? -1 : one;
    b->info.resize(num_scenes);
    PG3_HIP(hipSetDevice(ctx->device));
    PG3_HIP(b->views.Alloc(sizeof(SceneView) * num_scenes));
    PG3_HIP(hipMemcpyAsync(b->views.p, views.data(), sizeof(SceneView) * num_scenes, hipMemcpyHostToDevice,
                           ctx->stream));
    PG3_HIP(b->work.Alloc(sizeof(uint32_t) * (2 * kBatchParts + 2)));
    PG3_HIP(hipMemsetAsync(b->work.p, 0, sizeof(uint32_t) * (kBatchParts + 1), ctx->stream));
    if (int rc = BuildBatchItems(b.get())) return rc;
    *batch = b.release();
    return PG3G_OK;
}

int pg3g_batch_destroy(pg3g_batch *batch) {
    if (!batch) return PG3G_OK;
    (void)hipSetDevice(batch->ctx->device);
    (void)hipStreamSynchronize(batch->ctx->stream);
    delete batch;
    return PG3G_OK;
}

int pg3g_batch_set_config(pg3g_batch *batch, const pg3g_batch_config *config) {
    if (!batch || !config) return Fail(PG3G_ERR_INVALID_ARG, "null argument");
    if (config->iters_per_item == 0 || config->waves_per_cu < 4 || config->waves_per_cu % 4 != 0 ||
        config->waves_per_cu > 64)
        return Fail(PG3G_ERR_INVALID_ARG, "bad batch configuration");
    batch->cfg = *config;
    batch->cfg.use_graph = config->use_graph ? 1u : 0u;
    return BuildBatchItems(batch);
}

int pg3g_batch_get_config(const pg3g_batch *batch, pg3g_batch_config *config) {
    if (!batch || !config) return Fail(PG3G_ERR_INVALID_ARG, "null argument");
    *config = batch->cfg;
    return PG3G_OK;
}

namespace {
// Launches the iterations pg3g_run_iteration queued on the batch's scenes (they precede the pass).
int FlushBatchSlots(pg3g_batch *b) {
    for (uint32_t s : b->slots)
        if (int rc = FlushQueued(b->ctx, GetSlot(b->ctx, s))) return rc;
    return PG3G_OK;
}
}  // namespace

int pg3g_batch_render(pg3g_batch *batch, uint32_t passes) {
    if (int rc = BatchCheck(batch)) return rc;
    if (int rc = FlushBatchSlots(batch)) return rc;
    return BatchRenderPasses(batch, passes);
}

int pg3g_batch_render_timed(pg3g_batch *batch, uint32_t passes, float *elapsed_ms) {
    if (int rc = BatchCheck(batch)) return rc;
    if (int rc = FlushBatchSlots(batch)) return rc;
    return BatchTimed(batch, passes, elapsed_ms);
}

int pg3g_batch_autotune(pg3g_batch *batch, uint32_t max_candidates, float *best_pass_ms) {
    if (int rc = BatchCheck(batch)) return rc;
    // Work-item length (iterations of one tile) x persistent wavefronts per CU; the first
    // candidates are the ones most often best, so a small max_candidates still covers them.
    static const uint32_t kCand[][2] = {{2, 16}, {4, 16}, {1, 16}, {8, 16}, {16, 16}, {2, 8}, {4, 8}, {8, 32}};
    const uint32_t n = (max_candidates == 0) ? 8u : std::min<uint32_t>(max_candidates, 8u);
    const uint32_t useGraph = batch->cfg.use_graph;
    float best = 3.4e38f;
    pg3g_batch_config bestCfg = batch->cfg;
    for (uint32_t c = 0; c < n; ++c) {
        pg3g_batch_config cfg{kCand[c][0], kCand[c][1], useGraph};
        if (int rc = pg3g_batch_set_config(batch, &cfg)) return rc;
        float ms = 0.f;
        if (int rc = BatchTimed(batch, 1, &ms)) return rc;
        if (ms < best) { best = ms; bestCfg = cfg; }
    }
    if (int rc = pg3g_batch_set_config(batch, &bestCfg)) return rc;
    if (int rc = BatchRestart(batch)) return rc;
    if (best_pass_ms) *best_pass_ms = best;
    return PG3G_OK;
}

int pg3g_diff_setup(pg3g_context *ctx, uint32_t slot, uint32_t num_materials, const uint32_t *material_ids,
                    uint32_t *num_scalars) {
    SceneSlot *S = nullptr;
    if (int rc = CheckDiffSlot(ctx, slot, S)) return rc;
    if (num_materials > 0 && !material_ids) return Fail(PG3G_ERR_INVALID_ARG, "material_ids is null");
    std::vector<int32_t> base(S->hostMats.size(), -1);
    uint32_t n = 0;
    for (uint32_t i = 0; i < num_materials; ++i) {
        const uint32_t m = material_ids[i];
        if (m >= S->hostMats.size()) return Fail(PG3G_ERR_INVALID_ARG, "material id out of range");
        if (base[m] >= 0) return Fail(PG3G_ERR_INVALID_ARG, "material listed twice");
        const uint32_t c = DiffScalarCount(S->hostMats[m].type);
        if (c == 0)
            return Fail(PG3G_ERR_UNSUPPORTED, "differentiable materials: Lambert (albedo) and GGX conductor (tint, alpha)");
        base[m] = (int32_t)n;
        n += c;
    }
    if (n > kMaxDiffScalars) return Fail(PG3G_ERR_INVALID_ARG, "too many differentiable parameters");
    PG3_HIP(hipSetDevice(ctx->device));
    S->diffMats.assign(material_ids, material_ids + num_materials);
    S->gradBaseHost = base;
    S->numScalars = n;
    PG3_HIP(Upload(S->gradBase, base.data(), base.size(), ctx->stream));
    S->view.gradBase = S->gradBase.as<int32_t>();
    if (num_scalars) *num_scalars = n;
    return SyncView(ctx, *S);
}

int pg3g_diff_get_params(pg3g_context *ctx, uint32_t slot, float *scalars, float *em_texels) {
    SceneSlot *S = nullptr;
    if (int rc = CheckDiffSlot(ctx, slot, S)) return rc;
    if (scalars) {
        for (uint32_t m : S->diffMats) {
            const pg3g_material &d = S->evalHostMats[m];
            float *o = scalars + S->gradBaseHost[m];
            if (d.type == PG3G_MAT_LAMBERT) { o[0] = d.diffuse[0]; o[1] = d.diffuse[1]; o[2] = d.diffuse[2]; }
            else { o[0] = d.glossy[0]; o[1] = d.glossy[1]; o[2] = d.glossy[2]; o[3] = d.alpha; }
        }
    }
    if (em_texels) {
        if (S->view.em.w == 0) return Fail(PG3G_ERR_INVALID_ARG, "the scene has no EM image");
        const size_t n = (size_t)S->view.em.w * S->view.em.h;
        std::vector<float4> tex(n);
        PG3_HIP(hipSetDevice(ctx->device));
        PG3_HIP(hipMemcpyAsync(tex.data(), S->view.evalTexels, sizeof(float4) * n, hipMemcpyDeviceToHost, ctx->stream));
        PG3_HIP(hipStreamSynchronize(ctx->stream));
        for (size_t i = 0; i < n; ++i) { em_texels[i * 3] = tex[i].x; em_texels[i * 3 + 1] = tex[i].y; em_texels[i * 3 + 2] = tex[i].z; }
    }
    return PG3G_OK;
}

int pg3g_diff_set_params(pg3g_context *ctx, uint32_t slot, const float *scalars, const float *em_texels,
                         int refresh_sampling) {
    SceneSlot *S = nullptr;
    if (int rc = CheckDiffSlot(ctx, slot, S)) return rc;
    if (em_texels && S->view.em.w == 0) return Fail(PG3G_ERR_INVALID_ARG, "the scene has no EM image");
    if (int rc = FlushQueued(ctx, S)) return rc;                  // queued iterations use the old parameters
    PG3_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    SceneView &v = S->view;
    if (scalars) {
        for (uint32_t m : S->diffMats) {
            pg3g_material &d = S->evalHostMats[m];
            const float *p = scalars + S->gradBaseHost[m];
            for (int c = 0; c < 3; ++c) {
                if (!(p[c] >= 0.f)) return Fail(PG3G_ERR_INVALID_ARG, "albedo / tint must be >= 0");
                (d.type == PG3G_MAT_LAMBERT ? d.diffuse[c] : d.glossy[c]) = p[c];
            }
            if (d.type == PG3G_MAT_GGX_CONDUCTOR) {
                if (!(p[3] > 0.f)) return Fail(PG3G_ERR_INVALID_ARG, "alpha must be > 0");
                d.alpha = p[3];
            }
        }
    }
    std::vector<DevMaterial> em(S->evalHostMats.size());
    for (size_t i = 0; i < em.size(); ++i) em[i] = ToDev(S->evalHostMats[i]);
    PG3_HIP(Upload(S->evalMats, em.data(), em.size(), st));
    const size_t n = (size_t)v.em.w * v.em.h;
    if (em_texels) {
        std::vector<float4> tex(n);
        for (size_t i = 0; i < n; ++i) {
            if (!(em_texels[i * 3] >= 0.f && em_texels[i * 3 + 1] >= 0.f && em_texels[i * 3 + 2] >= 0.f))
                return Fail(PG3G_ERR_INVALID_ARG, "EM texels must be finite and >= 0");
            tex[i] = make_float4(em_texels[i * 3], em_texels[i * 3 + 1], em_texels[i * 3 + 2], 0.f);
        }
        PG3_HIP(Upload(S->evalTexels, tex.data(), n, st));
        S->evalHostTexels.assign(em_texels, em_texels + n * 3);
        S->evalTexDirty = true;
    } else if (v.em.w > 0 && !S->evalTexels.p) {
        PG3_HIP(S->evalTexels.Alloc(S->texels.bytes));
        PG3_HIP(hipMemcpyAsync(S->evalTexels.p, S->texels.p, S->texels.bytes, hipMemcpyDeviceToDevice, st));
    }
    if (refresh_sampling) {
        // the evaluation parameters become the sampling parameters
        S->hostMats = S->evalHostMats;
        PG3_HIP(Upload(S->mats, em.data(), em.size(), st));
        v.mats = S->mats.as<DevMaterial>();
        v.evalMats = v.mats;
        if (v.em.w > 0) {
            if (S->evalTexDirty) {
                S->hostTexels = S->evalHostTexels;
                PG3_HIP(hipMemcpyAsync(S->texels.p, S->evalTexels.p, sizeof(float4) * n, hipMemcpyDeviceToDevice, st));
                if (int rc = RebuildEmTables(ctx, *S)) return rc;
                S->evalTexDirty = false;
            }
            v.evalTexels = v.em.texels;
        }
    } else {
        v.evalMats = S->evalMats.as<DevMaterial>();
        if (v.em.w > 0) v.evalTexels = S->evalTexels.as<float4>();
    }
    return SyncView(ctx, *S);
}

}  // extern "C"

namespace {

// Differentiable forward of iterations [first, first + n) of one view (views == nullptr: the
// slot's camera, into its framebuffer) or of every view of pg3g_diff_set_views (into fbViews,
// cleared first).  Keeps the per-sample radiance for an adjoint of the same samples (up to 2^26
// samples).
int RenderDiffImpl(pg3g_context *ctx, SceneSlot *S, uint32_t first_iter, uint32_t n_iter, const DevView *views,
                   uint32_t nv, double *fb, float *elapsed_ms) {
    if ((uint64_t)first_iter + n_iter > 0xFFFFFFFFull) return Fail(PG3G_ERR_INVALID_ARG, "iteration range overflows");
    PG3_HIP(hipSetDevice(ctx->device));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (elapsed_ms) {
        PG3_HIP(hipEventCreate(&e0));
        PG3_HIP(hipEventCreate(&e1));
        PG3_HIP(hipEventRecord(e0, ctx->stream));
    }
    const uint64_t ns = (uint64_t)n_iter * nv * S->view.resX * S->view.resY;
    float *sl = nullptr;
    S->sampleValid = false;
    if (ns <= (1ull << 26)) {
        PG3_HIP(S->sampleL.Alloc(sizeof(float) * 3 * ns));
        sl = S->sampleL.as<float>();
    }
    hipError_t e = LaunchDiff(S->view, S->viewDev.as<SceneView>(), first_iter, n_iter, fb, nullptr, nullptr, nullptr,
                              0, sl, nullptr, ctx->numCUs, ctx->counters.as<unsigned long long>(), ctx->stream, views,
                              nv);
    if (elapsed_ms) {
        if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *elapsed_ms = ms;
    }
    PG3_HIP(e);
    if (sl) {
        S->sampleValid = true;
        S->sampleFirst = first_iter;
        S->sampleCount = n_iter;
        S->sampleEpoch = S->epoch;
        S->sampleViews = nv;
    }
    return PG3G_OK;
}

// Adjoint of iterations [first, first + n) of one view or of every view (gradients summed over
// the views): dloss holds dLoss/dImage of each view (nv images of res_x x res_y x 3).
int AdjointImpl(pg3g_context *ctx, SceneSlot *S, uint32_t first_iter, uint32_t n_iter, const float *dloss,
                const DevView *views, uint32_t nv, double *grad_scalars, double *grad_em, float *elapsed_ms) {
    if (!dloss) return Fail(PG3G_ERR_INVALID_ARG, "dloss_dimage is null");
    if (n_iter == 0) return Fail(PG3G_ERR_INVALID_ARG, "n_iter must be >= 1");
    if ((uint64_t)first_iter + n_iter > 0xFFFFFFFFull) return Fail(PG3G_ERR_INVALID_ARG, "iteration range overflows");
    const SceneView &v = S->view;
    PG3_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const size_t npx = (size_t)v.resX * v.resY;
    PG3_HIP(Upload(S->adjImg, dloss, npx * 3 * nv, st));
    PG3_HIP(S->gradScal.Alloc(sizeof(double) * kMaxDiffScalars));
    PG3_HIP(S->adjWork.Alloc(sizeof(uint32_t) * nv));                 // one item counter per view
    PG3_HIP(hipMemsetAsync(S->adjWork.p, 0, sizeof(uint32_t) * nv, st));
    PG3_HIP(hipMemsetAsync(S->gradScal.p, 0, sizeof(double) * kMaxDiffScalars, st));
    const size_t emN = (size_t)v.em.w * v.em.h * 3;
    double *gEm = nullptr;
    if (emN > 0) {
        PG3_HIP(S->gradEm.Alloc(sizeof(double) * emN));
        PG3_HIP(hipMemsetAsync(S->gradEm.p, 0, sizeof(double) * emN, st));
        gEm = S->gradEm.as<double>();
    }
    // radiance of the same samples kept by the last differentiable forward (same parameters and
    // views); without it a forward pass computes it first (cheaper than interleaving a primal
    // replay with the gradient replay in every wavefront)
    bool stored = S->sampleValid && S->sampleFirst == first_iter && S->sampleCount == n_iter &&
                  S->sampleEpoch == S->epoch && S->sampleViews == nv;
    const uint64_t ns = (uint64_t)n_iter * nv * npx;
    bool primal = false;
    if (!stored && ns <= (1ull << 26)) {
        PG3_HIP(S->sampleL.Alloc(sizeof(float) * 3 * ns));
        S->sampleValid = false;
        stored = primal = true;
    }
    hipEvent_t e0, e1;
    PG3_HIP(hipEventCreate(&e0));
    PG3_HIP(hipEventCreate(&e1));
    hipError_t e = hipEventRecord(e0, st);
    if (e == hipSuccess && primal)
        e = LaunchDiff(v, S->viewDev.as<SceneView>(), first_iter, n_iter, nullptr, nullptr, nullptr, nullptr, 0,
                       S->sampleL.as<float>(), nullptr, ctx->numCUs, ctx->counters.as<unsigned long long>(), st, views,
                       nv);
    if (e == hipSuccess)
        e = LaunchDiff(v, S->viewDev.as<SceneView>(), first_iter, n_iter, nullptr, S->adjImg.as<float>(), gEm,
                       S->gradScal.as<double>(), S->numScalars, stored ? S->sampleL.as<float>() : nullptr,
                       S->adjWork.as<uint32_t>(), ctx->numCUs, ctx->counters.as<unsigned long long>(), st, views, nv);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    PG3_HIP(e);
    if (primal) {
        S->sampleValid = true;
        S->sampleFirst = first_iter;
        S->sampleCount = n_iter;
        S->sampleEpoch = S->epoch;
        S->sampleViews = nv;
    }
    if (elapsed_ms) *elapsed_ms = ms;
    if (grad_scalars && S->numScalars > 0) {
        PG3_HIP(hipMemcpy(grad_scalars, S->gradScal.p, sizeof(double) * S->numScalars, hipMemcpyDeviceToHost));
        // alpha is clamped to [0.001, 1] by the material (materials.hxx:805): no gradient outside
        for (uint32_t m : S->diffMats)
            if (S->evalHostMats[m].type == PG3G_MAT_GGX_CONDUCTOR) {
                const float a = S->evalHostMats[m].alpha;
                if (a < 0.001f || a > 1.0f) grad_scalars[S->gradBaseHost[m] + 3] = 0.0;
            }
    }
    if (grad_em) {
        if (emN == 0) return Fail(PG3G_ERR_INVALID_ARG, "the scene has no EM image");
        PG3_HIP(hipMemcpy(grad_em, S->gradEm.p, sizeof(double) * emN, hipMemcpyDeviceToHost));
    }
    return PG3G_OK;
}

int CheckViews(SceneSlot *S) {
    if (S->numViews == 0) return Fail(PG3G_ERR_INVALID_ARG, "the slot has no views (pg3g_diff_set_views)");
    return PG3G_OK;
}

}  // namespace

extern "C" {

int pg3g_render_diff(pg3g_context *ctx, uint32_t slot, uint32_t first_iter, uint32_t n_iter, float *elapsed_ms) {
    SceneSlot *S = nullptr;
    if (int rc = CheckDiffSlot(ctx, slot, S)) return rc;
    if (int rc = FlushQueued(ctx, S)) return rc;
    if (int rc = RenderDiffImpl(ctx, S, first_iter, n_iter, nullptr, 1, S->fb.as<double>(), elapsed_ms)) return rc;
    S->iterations += n_iter;
    return PG3G_OK;
}

int pg3g_adjoint(pg3g_context *ctx, uint32_t slot, uint32_t first_iter, uint32_t n_iter, const float *dloss_dimage,
                 double *grad_scalars, double *grad_em, float *elapsed_ms) {
    SceneSlot *S = nullptr;
    if (int rc = CheckDiffSlot(ctx, slot, S)) return rc;
    return AdjointImpl(ctx, S, first_iter, n_iter, dloss_dimage, nullptr, 1, grad_scalars, grad_em, elapsed_ms);
}

int pg3g_diff_set_views(pg3g_context *ctx, uint32_t slot, uint32_t num_views, const pg3g_camera_desc *cameras,
                        const uint32_t *scene_ids) {
    SceneSlot *S = nullptr;
    if (int rc = CheckDiffSlot(ctx, slot, S)) return rc;
    if (num_views > PG3G_MAX_DIFF_VIEWS) return Fail(PG3G_ERR_INVALID_ARG, "too many views");
    if (num_views > 0 && (!cameras || !scene_ids)) return Fail(PG3G_ERR_INVALID_ARG, "cameras or scene_ids is null");
    std::vector<DevView> vs(num_views);
    for (uint32_t i = 0; i < num_views; ++i) {
        const pg3g_camera_desc &c = cameras[i];
        if (c.res_x != S->view.resX || c.res_y != S->view.resY)
            return Fail(PG3G_ERR_INVALID_ARG, "every view must have the slot's resolution");
        std::memset(&vs[i], 0, sizeof(DevView));
        CameraMatrix(c, vs[i].r2w);
        vs[i].camPos = mk(c.position[0], c.position[1], c.position[2]);
        vs[i].sceneId = scene_ids[i] & 0xFFFFu;
    }
    PG3_HIP(hipSetDevice(ctx->device));
    if (num_views > 0) {
        PG3_HIP(Upload(S->viewCams, vs.data(), vs.size(), ctx->stream));
        PG3_HIP(S->fbViews.Alloc(sizeof(double) * 3 * (size_t)S->view.resX * S->view.resY * num_views));
        PG3_HIP(hipStreamSynchronize(ctx->stream));
    }
    S->numViews = num_views;
    S->sampleValid = false;
    return PG3G_OK;
}

int pg3g_render_diff_views(pg3g_context *ctx, uint32_t slot, uint32_t first_iter, uint32_t n_iter, double *sums,
                           float *elapsed_ms) {
    SceneSlot *S = nullptr;
    if (int rc = CheckDiffSlot(ctx, slot, S)) return rc;
    if (int rc = CheckViews(S)) return rc;
    PG3_HIP(hipSetDevice(ctx->device));
    PG3_HIP(hipMemsetAsync(S->fbViews.p, 0, S->fbViews.bytes, ctx->stream));
    if (int rc = RenderDiffImpl(ctx, S, first_iter, n_iter, S->viewCams.as<DevView>(), S->numViews,
                                S->fbViews.as<double>(), elapsed_ms))
        return rc;
    if (sums) {
        PG3_HIP(hipMemcpyAsync(sums, S->fbViews.p, sizeof(double) * 3 * (size_t)S->view.resX * S->view.resY * S->numViews,
                               hipMemcpyDeviceToHost, ctx->stream));
        PG3_HIP(hipStreamSynchronize(ctx->stream));
    }
    return PG3G_OK;
}

int pg3g_adjoint_views(pg3g_context *ctx, uint32_t slot, uint32_t first_iter, uint32_t n_iter,
                       const float *dloss_dimages, double *grad_scalars, double *grad_em, float *elapsed_ms) {
    SceneSlot *S = nullptr;
    if (int rc = CheckDiffSlot(ctx, slot, S)) return rc;
    if (int rc = CheckViews(S)) return rc;
    return AdjointImpl(ctx, S, first_iter, n_iter, dloss_dimages, S->viewCams.as<DevView>(), S->numViews, grad_scalars,
                       grad_em, elapsed_ms);
}

int pg3g_adjoint_noalloc(pg3g_context *ctx, uint32_t slot, uint32_t first_iter, uint32_t n_iter,
                         const float *dloss_dimage_dev, double *grad_scalars_dev, double *grad_em_dev) {
    SceneSlot *S = nullptr;
    if (int rc = CheckDiffSlot(ctx, slot, S)) return rc;
    if (!dloss_dimage_dev) return Fail(PG3G_ERR_INVALID_ARG, "dloss_dimage_dev is null");
    if (n_iter == 0) return Fail(PG3G_ERR_INVALID_ARG, "n_iter must be >= 1");
    if ((uint64_t)first_iter + n_iter > 0xFFFFFFFFull) return Fail(PG3G_ERR_INVALID_ARG, "iteration range overflows");
    const SceneView &v = S->view;
    const uint64_t emN = (uint64_t)v.em.w * v.em.h * 3u;
    if (emN > kNoAllocMaxEmChannels)
        return Fail(PG3G_ERR_UNSUPPORTED, "pg3g_adjoint_noalloc holds EM gradients of at most 256x128 RGB texels "
                                          "(PG3G_NOALLOC_MAX_EM_CHANNELS); use pg3g_adjoint");
    if (grad_em_dev && emN == 0) return Fail(PG3G_ERR_INVALID_ARG, "the scene has no EM image");
    if (!ctx->noAllocReady) return Fail(PG3G_ERR_HIP, "the scratch array of pg3g_adjoint_noalloc is not available");
    // scalars with no gradient: GGX alpha outside the clamp range [0.001, 1] (materials.hxx:805)
    uint64_t zeroMask = 0;
    for (uint32_t m : S->diffMats)
        if (S->evalHostMats[m].type == PG3G_MAT_GGX_CONDUCTOR) {
            const float a = S->evalHostMats[m].alpha;
            if (a < 0.001f || a > 1.0f) zeroMask |= 1ull << (S->gradBaseHost[m] + 3);
        }
    const bool stored = S->sampleValid && S->sampleFirst == first_iter && S->sampleCount == n_iter &&
                        S->sampleEpoch == S->epoch && S->sampleViews == 1;
    PG3_HIP(hipSetDevice(ctx->device));
    PG3_HIP(LaunchAdjointNoAlloc(ctx->device, v, S->viewDev.as<SceneView>(), first_iter, n_iter, dloss_dimage_dev,
                                 S->numScalars, zeroMask, stored ? S->sampleL.as<float>() : nullptr, grad_scalars_dev,
                                 grad_em_dev, ctx->numCUs, ctx->counters.as<unsigned long long>(), ctx->stream));
    return PG3G_OK;
}

}  // extern "C"
