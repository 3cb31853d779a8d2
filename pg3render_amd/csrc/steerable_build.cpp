// Host build of the steerable EM sampler tables: icosahedron triangulation refined to the
// requested levels, vertex SH weights, and the 6-ary sum tree flattened level by level for the GPU.
// Follows EnvironmentMapSteerableSampler::TriangulateEm / RefineEmTriangulation /
// BuildTriangleTree (em_steerable_sampler.hxx:2963-3472, 3755-3793).  The tree layout is fixed by
// the reference's grouping: consecutive runs of at most 6 nodes of one layer form one node of the
// next layer, so node i of level l+1 owns nodes [6i, 6i+6) of level l and no pointers are needed.
//
// The reference refines depth first, one triangle at a time.  Whether a triangle is subdivided
// depends only on the triangle (its vertices' directions and luminances and its level), so the
// build here decides a whole level of triangles at once on worker threads, then walks the
// resulting subdivision tree in the reference's depth-first order to list the leaves.  The result
// is identical for any thread count.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <thread>

#include "pg3g_internal.hpp"

namespace pg3 {
namespace {

constexpr float kPi = 3.14159265358979f;
constexpr float kHalfPi = 1.5707963f;

struct Img {
    const float *t;
    uint32_t w, h;
    bool bilinear;
    F3 At(uint32_t x, uint32_t y) const {
        const float *p = t + ((size_t)w * y + x) * 3;
        return mk(p[0], p[1], p[2]);
    }
    void Wrap(int32_t x, int32_t y, uint32_t &ox, uint32_t &oy) const {   // em_image.hxx:245-278
        int32_t rx = x, ry = y;
        if (y >= (int32_t)h) { rx = x + (int32_t)(w / 2); ry = (int32_t)h - 1; }
        else if (y < 0) { rx = x + (int32_t)(w / 2); ry = 0; }
        int32_t m = rx % (int32_t)w;
        if (m < 0) m += (int32_t)w;
        ox = (uint32_t)m;
        oy = (uint32_t)ry;
    }
    F3 Eval(F3 d) const {                                                     // em_image.hxx:95-150
        float u, v;
        Dir2LatLong(d, u, v);
        const float xf = u * (float)w, yf = v * (float)h;
        if (!bilinear) return At(std::min((uint32_t)xf, w - 1u), std::min((uint32_t)yf, h - 1u));
        const int32_t cx = (int32_t)(xf + 0.5f), cy = (int32_t)(yf + 0.5f);
        const float xl = xf - ((float)(cx - 1) + 0.5f), yl = yf - ((float)(cy - 1) + 0.5f);
        uint32_t ax, ay, bx, by;
        Wrap(cx - 1, cy - 1, ax, ay);
        Wrap(cx, cy, bx, by);
        const F3 a = At(ax, ay) * (1 - xl) + At(bx, ay) * xl;
        const F3 b = At(ax, by) * (1 - xl) + At(bx, by) * xl;
        return a * (1 - yl) + b * yl;
    }
    float Lum(F3 d) const { return Luminance(Eval(d)); }
};

// A triangle of the refinement: vertex directions and the EM luminance at them (CreateNewVertex,
// :3063-3085, evaluates it once per vertex).
struct Tri {
    F3 d[3];
    float lum[3];
    uint32_t level;
};

struct Builder {
    const Img &img;
    float maxErr;
    uint32_t minLevel, maxLevel;
    // BuildParameters debug defaults (em_steerable_sampler.hxx:44-56)
    const uint32_t maxSamplesPerDim = 50;
    const float oversampling = 0.7f, maxSpan = 1.1f;

    static void Weight(F3 d, float f, float out[9]) {                 // GenerateSphHarm(dir, lum), :181-202
        out[0] = f * 0.282095f;
        out[1] = f * 0.488603f * d.y;
        out[2] = f * 0.488603f * d.z;
        out[3] = f * 0.488603f * d.x;
        out[4] = f * 1.092548f * d.x * d.y;
        out[5] = f * 1.092548f * d.y * d.z;
        out[6] = f * 0.315392f * (3.f * d.z * d.z - 1.f);
        out[7] = f * 1.092548f * d.x * d.z;
        out[8] = f * 0.546274f * (d.x * d.x - d.y * d.y);
    }
    // Linear luminance approximation over the whole triangle, :1445-1473
    static float Approx(const Tri &t, float bx, float by) {
        const float bz = std::min(std::max(1.0f - bx - by, 0.0f), 1.0f);
        return bx * t.lum[0] + by * t.lum[1] + bz * t.lum[2];
    }
    bool ErrorTooLarge(const Tri &whole, F3 s0, F3 s1, F3 s2, uint32_t spd) const {   // :3219-3274
        const float bin = 1.f / spd;
        const F3 w0 = whole.d[0], w1 = whole.d[1], w2 = whole.d[2];
        const F3 N = Cross(w1 - w0, w2 - w0);
        const float den = Dot(N, N);
        for (uint32_t i = 0; i <= spd; ++i) {
            for (uint32_t j = 0; j <= spd; ++j) {
                const float ux = (i * bin) * (i * bin), uy = j * bin;
                const float xs = std::sqrt(ux);
                const float b0 = 1.f - xs, b1 = uy * xs, b2 = 1.0f - b0 - b1;
                const F3 p = s0 * b0 + s1 * b1 + s2 * b2;
                // Geom::Triangle::BarycentricCoords, geom.hxx:634-674
                float bx = Dot(N, Cross(w2 - w1, p - w1)), by = Dot(N, Cross(w0 - w2, p - w2));
                bx /= den;
                by /= den;
                const float approx = Approx(whole, std::min(std::max(bx, 0.f), 1.f), std::min(std::max(by, 0.f), 1.f));
                const float em = img.Lum(Normalize(p));
                const float lo = std::min(em, approx), hi = std::max(em, approx);
                if ((hi + 0.0001f) > ((lo + 0.0001f) * (1.f + maxErr))) return true;
            }
        }
        return false;
    }
    bool NeedsSubdivImpl(F3 v0, float s0, F3 v1, float s1, F3 v2, float s2, const Tri &whole) const {  // :3276-3397
        if (img.h == 0 || img.w == 0) return false;
        const F3 cen = (v0 + v1 + v2) / 3.f;
        const F3 e01 = Normalize((v0 + v1) / 2.f), e12 = Normalize((v1 + v2) / 2.f), e20 = Normalize((v2 + v0) / 2.f);
        const F3 cd = Normalize(cen);
        const float se01 = std::sqrt(1.f - e01.z * e01.z), se12 = std::sqrt(1.f - e12.z * e12.z);
        const float se20 = std::sqrt(1.f - e20.z * e20.z), sc = std::sqrt(1.f - cd.z * cd.z);
        const float mn = std::min({s0, s1, s2, se01, se12, se20, sc});
        const float mx = std::max({s0, s1, s2, se01, se12, se20, sc});
        const float poleSin = std::sin(0.5f * kHalfPi / img.h);
        const float minSin = std::max(mn, poleSin), maxSin = std::max(mx, poleSin);
        // SubdivTestSamplesPerDim, :3160-3217
        const float pixTheta = kPi / img.h;
        const float a0 = std::min(std::min(pixTheta, minSin * (2.f * kPi) / img.w) / 2.0f, kHalfPi - 0.1f);
        const float a1 = std::min(std::min(pixTheta, maxSin * (2.f * kPi) / img.w) / 2.0f, kHalfPi - 0.1f);
        const float dist = Length(cen);
        const float ps0 = std::tan(a0) * dist, ps1 = std::tan(a1) * dist;
        const float avg = (LenSqr(v0 - v1) + LenSqr(v1 - v2) + LenSqr(v2 - v0)) / 3.0f;
        const float maxSpd = std::sqrt((avg / ((ps0 * ps0) / 2.0f)) / 2.f) * oversampling;
        const float minSpd = std::sqrt((avg / ((ps1 * ps1) / 2.0f)) / 2.f) * oversampling;
        const uint32_t maxSpdI = (uint32_t)std::ceil(maxSpd);
        if (maxSamplesPerDim != 0 && maxSamplesPerDim < maxSpdI) return true;
        if ((maxSpd / minSpd) >= maxSpan && maxSpd > 32.0f) {
            return NeedsSubdivImpl(v0, s0, e01, se01, e20, se20, whole) ||
                   NeedsSubdivImpl(v1, s1, e12, se12, e01, se01, whole) ||
                   NeedsSubdivImpl(v2, s2, e20, se20, e12, se12, whole) ||
                   NeedsSubdivImpl(e01, se01, e12, se12, e20, se20, whole);
        }
        return ErrorTooLarge(whole, v0, v1, v2, maxSpdI);
    }
    bool NeedsSubdiv(const Tri &t) const {                                   // :3399-3433
        if (t.level < minLevel) return true;
        if (t.level >= maxLevel) return false;
        const F3 d0 = t.d[0], d1 = t.d[1], d2 = t.d[2];
        return NeedsSubdivImpl(d0, std::sqrt(1.f - d0.z * d0.z), d1, std::sqrt(1.f - d1.z * d1.z), d2,
                               std::sqrt(1.f - d2.z * d2.z), t);
    }
    // SubdivideTriangle, :3436-3472: children in the order the reference refines them (it pushes
    // the centre, then the corners at v0, v1, v2 to the front of its queue)
    void Children(const Tri &c, Tri out[4]) const {
        const F3 n0 = Normalize((c.d[0] + c.d[1]) / 2.f);
        const F3 n1 = Normalize((c.d[1] + c.d[2]) / 2.f);
        const F3 n2 = Normalize((c.d[2] + c.d[0]) / 2.f);
        const float l0 = img.Lum(n0), l1 = img.Lum(n1), l2 = img.Lum(n2);
        const uint32_t L = c.level + 1;
        out[0] = Tri{{n1, c.d[2], n2}, {l1, c.lum[2], l2}, L};
        out[1] = Tri{{n0, c.d[1], n1}, {l0, c.lum[1], l1}, L};
        out[2] = Tri{{c.d[0], n0, n2}, {c.lum[0], l0, l2}, L};
        out[3] = Tri{{n0, n1, n2}, {l0, l1, l2}, L};
    }
};

// Runs fn(i) for i in [0, n) on `threads` threads (blocks of indices from a shared counter).
template <class Fn>
void ParallelFor(size_t n, uint32_t threads, const Fn &fn) {
    const size_t block = 8;
    const uint32_t t = (uint32_t)std::min<size_t>(threads, (n + block - 1) / block);
    if (t <= 1) {
        for (size_t i = 0; i < n; ++i) fn(i);
        return;
    }
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (;;) {
            const size_t b = next.fetch_add(block);
            if (b >= n) return;
            for (size_t i = b; i < std::min(n, b + block); ++i) fn(i);
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t k = 1; k < t; ++k) pool.emplace_back(work);
    work();
    for (std::thread &th : pool) th.join();
}

}  // namespace

uint32_t SteerableBuildThreads() {
    const char *names[] = {"PG3G_BUILD_THREADS", "OMP_NUM_THREADS"};
    for (const char *nm : names)
        if (const char *e = std::getenv(nm)) {
            const long v = std::strtol(e, nullptr, 10);
            if (v > 0) return (uint32_t)std::min<long>(v, 256);
        }
    return std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
}

bool BuildSteerable(const float *texels, uint32_t w, uint32_t h, bool bilinear, float maxErr, uint32_t minLevel,
                    uint32_t maxLevel, SteerableTables &out, std::string &err, uint32_t threads) {
    if (threads == 0) threads = SteerableBuildThreads();
    if (!texels || w == 0 || h == 0) { err = "steerable build: empty EM image"; return false; }
    if (maxLevel > 12) { err = "steerable build: max subdivision level > 12 is not supported"; return false; }
    const Img img{texels, w, h, bilinear};
    const Builder B{img, maxErr > 0.f ? maxErr : 5.f, minLevel, maxLevel, 50, 0.7f, 1.1f};

    // Unit icosahedron, geom.hxx:291-346
    const float t = (1.0f + std::sqrt(5.0f)) / 2.0f;
    const float si = 1.0f / std::sqrt(1.0f + t * t);
    const float V[12][3] = {{t, 1, 0}, {-t, 1, 0}, {t, -1, 0}, {-t, -1, 0}, {1, 0, t}, {1, 0, -t},
                            {-1, 0, t}, {-1, 0, -t}, {0, t, 1}, {0, -t, 1}, {0, t, -1}, {0, -t, -1}};
    const uint32_t F[20][3] = {{0, 8, 4}, {1, 10, 7}, {2, 9, 11}, {7, 3, 1}, {0, 5, 10}, {3, 9, 6}, {3, 11, 9},
                               {8, 6, 4}, {2, 4, 9}, {3, 7, 11}, {4, 2, 0}, {9, 4, 6}, {2, 11, 5}, {0, 10, 8},
                               {5, 0, 2}, {10, 5, 7}, {1, 6, 8}, {1, 8, 10}, {6, 1, 3}, {11, 7, 5}};
    F3 base[12];
    float baseLum[12];
    for (int i = 0; i < 12; ++i) {
        base[i] = mk(V[i][0], V[i][1], V[i][2]) * si;
        baseLum[i] = img.Lum(base[i]);
    }

    // Refinement decided one level at a time: tri[l] holds the triangles of level l, first[l][i]
    // the index in tri[l+1] of the first of triangle i's four children (-1: a leaf).
    std::vector<std::vector<Tri>> tri(1);
    std::vector<std::vector<int64_t>> first;
    for (int i = 0; i < 20; ++i) {
        const uint32_t *f = F[i];
        tri[0].push_back(Tri{{base[f[0]], base[f[1]], base[f[2]]}, {baseLum[f[0]], baseLum[f[1]], baseLum[f[2]]}, 0});
    }
    uint64_t subdivided = 0;
    for (size_t l = 0; !tri[l].empty(); ++l) {
        const std::vector<Tri> &cur = tri[l];
        std::vector<uint8_t> sub(cur.size());
        ParallelFor(cur.size(), threads, [&](size_t i) { sub[i] = B.NeedsSubdiv(cur[i]) ? 1 : 0; });
        std::vector<int64_t> fc(cur.size(), -1);
        int64_t m = 0;
        for (size_t i = 0; i < cur.size(); ++i)
            if (sub[i]) { fc[i] = m; m += 4; }
        subdivided += (uint64_t)m / 4u;
        std::vector<Tri> next((size_t)m);
        ParallelFor(cur.size(), threads, [&](size_t i) { if (fc[i] >= 0) B.Children(cur[i], &next[(size_t)fc[i]]); });
        first.push_back(std::move(fc));
        tri.push_back(std::move(next));
    }
    // RefineEmTriangulation (:3090-3155) processes the queue from the front, pushes children to the
    // front and prepends accepted triangles to its output: the leaves in reverse depth-first order.
    std::vector<const Tri *> leaves;
    {
        std::vector<std::pair<uint32_t, int64_t>> stack;   // (level, index), visited last-in first-out
        for (int i = 19; i >= 0; --i) stack.emplace_back(0u, (int64_t)i);
        while (!stack.empty()) {
            const auto [l, i] = stack.back();
            stack.pop_back();
            const int64_t fc = first[l][(size_t)i];
            if (fc < 0) {
                leaves.push_back(&tri[l][(size_t)i]);
                continue;
            }
            for (int k = 3; k >= 0; --k) stack.emplace_back(l + 1, fc + k);
        }
        std::reverse(leaves.begin(), leaves.end());
    }

    // Level 0: leaf weights = area(planar) * mean vertex weight (:1274-1301)
    out.numLeaves = (uint32_t)leaves.size();
    out.numVertices = (uint32_t)(12u + 3u * subdivided);   // every subdivision creates 3 vertices
    out.leafVerts.assign((size_t)out.numLeaves * 36, 0.f);
    std::vector<float> level((size_t)out.numLeaves * 12, 0.f);
    ParallelFor(out.numLeaves, threads, [&](size_t i) {
        const Tri &tr = *leaves[i];
        float wv[3][9];
        for (int k = 0; k < 3; ++k) {
            Builder::Weight(tr.d[k], tr.lum[k], wv[k]);
            const F3 d = tr.d[k];
            float *lv = &out.leafVerts[i * 36 + k * 12];
            lv[0] = d.x; lv[1] = d.y; lv[2] = d.z;
            for (int j = 0; j < 9; ++j) lv[3 + j] = wv[k][j];
        }
        const F3 a = tr.d[0], b = tr.d[1], c = tr.d[2];
        const float area = Length(Cross(b - a, c - b)) / 2.f;                 // geom.hxx:253-284
        for (int j = 0; j < 9; ++j) level[i * 12 + j] = ((wv[0][j] + wv[1][j] + wv[2][j]) / 3.f) * area;
    });
    // Upper levels: sums of <= 6 consecutive children, accumulated from zero in child order
    out.levelCount.clear();
    out.levelBase.clear();
    out.nodeW.clear();
    uint32_t count = out.numLeaves;
    for (;;) {
        out.levelBase.push_back((uint32_t)(out.nodeW.size() / 12));
        out.levelCount.push_back(count);
        out.nodeW.insert(out.nodeW.end(), level.begin(), level.end());
        if (count <= 1) break;
        const uint32_t next = (count + 5) / 6;
        std::vector<float> up((size_t)next * 12, 0.f);
        for (uint32_t i = 0; i < next; ++i)
            for (uint32_t c = i * 6; c < std::min(count, i * 6 + 6); ++c)
                for (int j = 0; j < 9; ++j) up[(size_t)i * 12 + j] += level[(size_t)c * 12 + j];
        level.swap(up);
        count = next;
    }
    out.numLevels = (uint32_t)out.levelCount.size();
    if (out.numLevels > (uint32_t)kMaxTreeLevels) { err = "steerable build: tree too deep"; return false; }
    // Root weight: TriangleSetNode::ComputeWeight of the root = the single top-level node
    const float *root = &out.nodeW[(size_t)out.levelBase.back() * 12];
    for (int j = 0; j < 9; ++j) out.rootW[j] = root[j];
    return true;
}

}  // namespace pg3
