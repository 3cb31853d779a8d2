// Oracle (test infrastructure): environment-map image, 1-D/2-D distributions and the three EM
// samplers of the reference, restated on the CPU.
#pragma once

#include <array>
#include <deque>
#include <functional>
#include <list>
#include <memory>
#include <vector>

#include "orc_core.hpp"

namespace orc {

// ---------------------------------------------------------------------------------------------
// EnvironmentMapImage, em_image.hxx:16-288 (data built in memory; the EXR loader is not used)
struct EmImage {
    uint32_t w = 0, h = 0;
    bool bilinear = true;
    std::vector<V3> data;   // row-major w*y+x, em_image.hxx:193-194

    const V3 &At(uint32_t x, uint32_t y) const { return data[(size_t)w * y + x]; }
    // em_image.hxx:245-278
    void NormCoords(int32_t x, int32_t y, uint32_t &ox, uint32_t &oy) const {
        int32_t rx, ry;
        if (y >= (int32_t)h) { rx = x + (int32_t)(w / 2); ry = (int32_t)h - 1; }
        else if (y < 0) { rx = x + (int32_t)(w / 2); ry = 0; }
        else { rx = x; ry = y; }
        int32_t m = rx % (int32_t)w;
        if (m < 0) m += (int32_t)w;   // Math::ModX, math.hxx:141-150
        ox = (uint32_t)m; oy = (uint32_t)ry;
    }
    // em_image.hxx:95-142
    V3 Evaluate(float u, float v) const {
        const float xf = u * (float)w, yf = v * (float)h;
        if (!bilinear) {
            const uint32_t x0 = std::min((uint32_t)xf, w - 1u);
            const uint32_t y0 = std::min((uint32_t)yf, h - 1u);
            return At(x0, y0);
        }
        const int32_t cx = (int32_t)(xf + 0.5f), cy = (int32_t)(yf + 0.5f);
        const int32_t x0 = cx - 1, y0 = cy - 1;
        const float xl = xf - ((float)x0 + 0.5f);
        const float yl = yf - ((float)y0 + 0.5f);
        uint32_t ax, ay, bx, by;
        NormCoords(x0, y0, ax, ay);
        NormCoords(cx, cy, bx, by);
        // Filter::Triangle -> Math::Bilerp, filter.hxx:12-29, math.hxx:204-229
        const V3 a = At(ax, ay) * (1 - xl) + At(bx, ay) * xl;
        const V3 b = At(ax, by) * (1 - xl) + At(bx, by) * xl;
        return a * (1 - yl) + b * yl;
    }
    V3 Evaluate(const V3 &dir) const { const V2 uv = Dir2LatLong(dir); return Evaluate(uv.x, uv.y); }
    // em_image.hxx:152-183 with Filter::TriangleIntegral, filter.hxx:35-60
    float AvgPixelLum(uint32_t x, uint32_t y) const {
        if (!bilinear) return At(x, y).Luminance();
        uint32_t x0, y0, x2, y2;
        NormCoords((int32_t)x - 1, (int32_t)y - 1, x0, y0);
        NormCoords((int32_t)x + 1, (int32_t)y + 1, x2, y2);
        const uint32_t x1 = x, y1 = y;
        const float e = 1 / 24.f, c = 2 / 3.f;
        return At(x0, y0).Luminance() * e + At(x0, y1).Luminance() * e + At(x0, y2).Luminance() * e +
               At(x1, y0).Luminance() * e + At(x1, y1).Luminance() * c + At(x1, y2).Luminance() * e +
               At(x2, y0).Luminance() * e + At(x2, y1).Luminance() * e + At(x2, y2).Luminance() * e;
    }
};

// ---------------------------------------------------------------------------------------------
// Distribution1D (distribution.hxx:45-74 ComputeCdf, :391-455 SampleContinuous).  The hierarchical
// variant active in hard_config.hxx:20 selects the same segment as std::upper_bound over the full
// CDF, so the full CDF is searched directly.  The prefix sum is accumulated in FP64 (the reference
// accumulates in FP32); ComputeCdfF32 keeps the reference's FP32 accumulation for comparison.
struct Dist1D {
    std::vector<float> cdf;   // count+1
    float integral = 0.f;
    size_t count = 0;

    static void ComputeCdfF32(std::vector<float> &cdf, float &integral, const float *f, size_t n) {
        cdf.assign(n + 1, 0.f);
        for (size_t i = 1; i < n + 1; ++i) cdf[i] = cdf[i - 1] + f[i - 1] / n;
        integral = cdf[n];
        if (integral == 0.f) for (size_t i = 1; i < n + 1; ++i) cdf[i] = (float)i / n;
        else for (size_t i = 1; i < n + 1; ++i) cdf[i] /= integral;
    }
    void Build(const float *f, size_t n) {
        count = n;
        cdf.assign(n + 1, 0.f);
        std::vector<double> acc(n + 1, 0.0);
        for (size_t i = 1; i < n + 1; ++i) acc[i] = acc[i - 1] + (double)f[i - 1] / (double)n;
        integral = (float)acc[n];
        if (acc[n] == 0.0) for (size_t i = 1; i < n + 1; ++i) cdf[i] = (float)i / n;
        else for (size_t i = 1; i < n + 1; ++i) cdf[i] = (float)(acc[i] / acc[n]);
    }
    void Sample(float u, float &x, size_t &seg, float &pdf) const {
        const float ut = u * 0.999999f;
        const float *it = std::upper_bound(cdf.data(), cdf.data() + count + 1, ut);
        long s = (long)(it - cdf.data()) - 1;
        s = std::max(0L, std::min(s, (long)count - 1));
        seg = (size_t)s;
        const float p = cdf[seg + 1] - cdf[seg];
        const float off = (ut - cdf[seg]) / p;
        pdf = p * count;
        x = (seg + off) / count;
    }
    float Pdf(size_t seg) const { return (cdf[seg + 1] - cdf[seg]) * count; }
};

struct Dist2D {      // distribution.hxx:484-541
    std::vector<Dist1D> cond;
    Dist1D marg;
    void Build(const float *f, int nu, int nv) {
        cond.resize(nv);
        std::vector<float> mf(nv);
        for (int v = 0; v < nv; ++v) { cond[v].Build(f + (size_t)v * nu, nu); mf[v] = cond[v].integral; }
        marg.Build(mf.data(), nv);
    }
    void Sample(float r0, float r1, float &u, float &v, uint32_t &sx, uint32_t &sy, float &pdf) const {
        float mp, cp; size_t x, y;
        marg.Sample(r0, v, y, mp);
        cond[y].Sample(r1, u, x, cp);
        sx = (uint32_t)x; sy = (uint32_t)y;
        pdf = mp * cp;
    }
    float Pdf(float u, float v) const {
        const size_t nu = cond[0].count, nv = marg.count;
        size_t iu = std::min((size_t)std::max(0.f, u * nu), nu - 1);
        size_t iv = std::min((size_t)std::max(0.f, v * nv), nv - 1);
        return cond[iv].Pdf(iu) * marg.Pdf(iv);
    }
};

// EnvironmentMapSimpleSphericalSampler, em_simple_spherical_sampler.hxx:7-209
struct SimpleSampler {
    Dist2D dist;
    const EmImage *img = nullptr;
    const float plan2ang = 1.0f / (2.0f * kPiF * kPiF);
    static float SinMidTheta(uint32_t h, uint32_t sy) { return sinf(kPiF * (sy + 0.5f) / h); }
    void Build(const EmImage &im) {
        img = &im;
        std::vector<float> src((size_t)im.w * im.h);
        for (uint32_t row = 0; row < im.h; ++row) {
            const float s = SinMidTheta(im.h, row);
            for (uint32_t col = 0; col < im.w; ++col)
                src[(size_t)row * im.w + col] = s * im.AvgPixelLum(col, row);
        }
        dist.Build(src.data(), (int)im.w, (int)im.h);
    }
    bool Sample(const Frame &fr, bool front, bool back, float u0, float u1, V3 &dir, float &pdf,
                V3 &radCos) const {
        float u, v, dp; uint32_t sx, sy;
        dist.Sample(u0, u1, u, v, sx, sy, dp);
        dir = LatLong2Dir(u, v);
        pdf = dp * plan2ang / SinMidTheta(img->h, sy);
        const V3 rad = !img->bilinear ? img->At(sx, sy) : img->Evaluate(u, v);
        const float c = Dot(dir, fr.Normal());
        if ((front && c > 0.0f) || (back && c < 0.0f)) radCos = rad * std::abs(c);
        else radCos = V3(0.f);
        return true;
    }
    float Pdf(const V3 &dir) const {   // ignores the frame and side flags (:84-98)
        const V2 uv = Dir2LatLong(dir);
        const uint32_t seg = std::min((uint32_t)(uv.y * img->h), img->h - 1u);
        return dist.Pdf(uv.x, uv.y) * plan2ang / SinMidTheta(img->h, seg);
    }
};

// ---------------------------------------------------------------------------------------------
// Steerable sampler (em_steerable_sampler.hxx)
typedef std::array<float, 9> SH9;
inline float Dot9(const SH9 &a, const SH9 &b) {        // :135-146
    float r = 0.0f;
    for (size_t i = 0; i < 9; i++) r += a[i] * b[i];
    return r;
}
inline SH9 ShBasis(const V3 &d, float f = 1.f) {        // GenerateSphHarm, :181-202
    SH9 b;
    b[0] = f * 0.282095f;
    b[1] = f * 0.488603f * d.y;
    b[2] = f * 0.488603f * d.z;
    b[3] = f * 0.488603f * d.x;
    b[4] = f * 1.092548f * d.x * d.y;
    b[5] = f * 1.092548f * d.y * d.z;
    b[6] = f * 0.315392f * (3.f * d.z * d.z - 1.f);
    b[7] = f * 1.092548f * d.x * d.z;
    b[8] = f * 0.546274f * (d.x * d.x - d.y * d.y);
    return b;
}
inline SH9 ClampedCosCoeffs(const V3 &n, bool comp) {   // GenerateForClampedCos, :657-688
    const float c1 = .429043f, c2 = .511644f, c3 = .743125f, c4 = .886227f, c5 = .247708f;
    SH9 c;
    c[0] = c4 + (comp ? 0.15f : 0.f);
    c[1] = 2 * c2 * n.y;
    c[2] = 2 * c2 * n.z;
    c[3] = 2 * c2 * n.x;
    c[4] = 2 * c1 * n.x * n.y;
    c[5] = 2 * c1 * n.y * n.z;
    c[6] = c3 * Sqr(n.z) - c5;
    c[7] = 2 * c1 * n.x * n.z;
    c[8] = c1 * (Sqr(n.x) - Sqr(n.y));
    return c;
}

// geom.hxx:291-346
inline void UnitIcosahedron(V3 v[12], uint32_t f[20][3]) {
    const float t = (1.0f + std::sqrt(5.0f)) / 2.0f;
    const float s = std::sqrt(1.0f + t * t);
    const float si = 1.0f / s;
    const float vv[12][3] = {{t, 1, 0}, {-t, 1, 0}, {t, -1, 0}, {-t, -1, 0}, {1, 0, t}, {1, 0, -t},
                             {-1, 0, t}, {-1, 0, -t}, {0, t, 1}, {0, -t, 1}, {0, t, -1}, {0, -t, -1}};
    for (int i = 0; i < 12; ++i) v[i] = V3(vv[i][0], vv[i][1], vv[i][2]) * si;
    const uint32_t ff[20][3] = {{0, 8, 4}, {1, 10, 7}, {2, 9, 11}, {7, 3, 1}, {0, 5, 10},
                                {3, 9, 6}, {3, 11, 9}, {8, 6, 4}, {2, 4, 9}, {3, 7, 11},
                                {4, 2, 0}, {9, 4, 6}, {2, 11, 5}, {0, 10, 8}, {5, 0, 2},
                                {10, 5, 7}, {1, 6, 8}, {1, 8, 10}, {6, 1, 3}, {11, 7, 5}};
    std::memcpy(f, ff, sizeof(ff));
}

inline float TriSurfaceArea(const V3 &a, const V3 &b, const V3 &c) {     // geom.hxx:253-284
    return Cross(b - a, c - b).Length() / 2.f;
}
// geom.hxx:634-674
inline V2 BarycentricCoords(const V3 &p, const V3 &v0, const V3 &v1, const V3 &v2) {
    const V3 N = Cross(v1 - v0, v2 - v0);
    const float den = Dot(N, N);
    V2 b;
    b.x = Dot(N, Cross(v2 - v1, p - v1));
    b.y = Dot(N, Cross(v0 - v2, p - v2));
    b.x /= den; b.y /= den;
    return b;
}

struct SteerableSampler {
    struct Vertex { V3 dir; SH9 w; };
    struct Tri { uint32_t vi[3]; uint32_t level; };
    struct Node {                 // TriangleSetNode or TriangleNode
        bool leaf = false;
        Tri tri{};
        std::vector<SH9> childW;  // inner nodes only (<= 6)
        std::vector<std::unique_ptr<Node>> child;
    };
    const EmImage *img = nullptr;
    float maxErr = 5.f;
    uint32_t minLevel = 4, maxLevel = 9;
    uint32_t maxSamplesPerDim = 50;     // BuildParameters debug defaults, :44-56
    float oversampling = 0.7f, maxSpan = 1.1f;
    std::vector<Vertex> verts;
    std::unique_ptr<Node> root;

    uint32_t NewVertex(const V3 &dir) {              // CreateNewVertex, :3063-3085
        const float lum = img->Evaluate(dir).Luminance();
        verts.push_back(Vertex{dir, ShBasis(dir, lum)});
        return (uint32_t)verts.size() - 1;
    }
    SH9 TriWeight(const Tri &t) const {                // TriangleNode::ComputeWeight, :1274-1301
        const Vertex &a = verts[t.vi[0]], &b = verts[t.vi[1]], &c = verts[t.vi[2]];
        const float area = TriSurfaceArea(a.dir, b.dir, c.dir);
        SH9 r;
        for (int i = 0; i < 9; ++i) r[i] = ((a.w[i] + b.w[i] + c.w[i]) / 3.f) * area;
        return r;
    }
    static SH9 NodeWeight(const Node &n) {            // TriangleSetNodeBase::ComputeWeight, :1167-1173
        SH9 r; r.fill(0.f);
        for (const auto &w : n.childW) for (int i = 0; i < 9; ++i) r[i] += w[i];
        return r;
    }

    // TriangleNode::EvaluateLuminanceApprox, :1445-1473
    float LumApprox(const Tri &t, float bx, float by) const {
        const float w = Clamp(1.0f - bx - by, 0.0f, 1.0f);
        const float l0 = img->Evaluate(verts[t.vi[0]].dir).Luminance();
        const float l1 = img->Evaluate(verts[t.vi[1]].dir).Luminance();
        const float l2 = img->Evaluate(verts[t.vi[2]].dir).Luminance();
        return bx * l0 + by * l1 + w * l2;
    }
    // IsEstimationErrorTooLarge, :3219-3274
    bool ErrorTooLarge(const Tri &whole, const V3 &s0, const V3 &s1, const V3 &s2, uint32_t spd) const {
        const float bin = 1.f / spd;
        for (uint32_t i = 0; i <= spd; ++i)
            for (uint32_t j = 0; j <= spd; ++j) {
                const V3 p = SampleUniformTriangle(s0, s1, s2, Sqr(i * bin), j * bin);
                const V2 b = BarycentricCoords(p, verts[whole.vi[0]].dir, verts[whole.vi[1]].dir,
                                               verts[whole.vi[2]].dir);
                const float approx = LumApprox(whole, Clamp(b.x, 0.f, 1.f), Clamp(b.y, 0.f, 1.f));
                const float em = img->Evaluate(Normalize(p)).Luminance();
                const float mn = std::min(em, approx), mx = std::max(em, approx);
                if ((mx + 0.0001f) > ((mn + 0.0001f) * (1.f + maxErr))) return true;
            }
        return false;
    }
    // SubdivTestSamplesPerDim, :3160-3217
    void SamplesPerDim(const V3 &v0, const V3 &v1, const V3 &v2, const V3 &centroid, float minSin,
                       float maxSin, float &minSpd, float &maxSpd) const {
        const float ph = kPiF / img->h;
        const float minA0 = std::min(ph, minSin * (2.f * kPiF) / img->w);
        const float maxA0 = std::min(ph, maxSin * (2.f * kPiF) / img->w);
        const float ub0 = std::min(minA0 / 2.0f, kPiDiv2F - 0.1f);
        const float ub1 = std::min(maxA0 / 2.0f, kPiDiv2F - 0.1f);
        const float dist = centroid.Length();
        const float ps0 = std::tan(ub0) * dist, ps1 = std::tan(ub1) * dist;
        const float e0 = (v0 - v1).LenSqr(), e1 = (v1 - v2).LenSqr(), e2 = (v2 - v0).LenSqr();
        const float avg = (e0 + e1 + e2) / 3.0f;
        const float g0 = Sqr(ps0) / 2.0f, g1 = Sqr(ps1) / 2.0f;
        float spd0 = std::sqrt((avg / g0) / 2.f), spd1 = std::sqrt((avg / g1) / 2.f);
        spd0 *= oversampling; spd1 *= oversampling;
        maxSpd = spd0;
        minSpd = spd1;
    }
    // TriangleHasToBeSubdividedImpl, :3276-3397
    bool NeedsSubdivImpl(const V3 &v0, float s0, const V3 &v1, float s1, const V3 &v2, float s2,
                         const Tri &whole) const {
        if (img->h == 0 || img->w == 0) return false;
        const V3 cen = (v0 + v1 + v2) / 3.f;
        const V3 e01 = Normalize((v0 + v1) / 2.f), e12 = Normalize((v1 + v2) / 2.f),
                 e20 = Normalize((v2 + v0) / 2.f), cd = Normalize(cen);
        const float se01 = std::sqrt(1.f - Sqr(e01.z)), se12 = std::sqrt(1.f - Sqr(e12.z)),
                    se20 = std::sqrt(1.f - Sqr(e20.z)), sc = std::sqrt(1.f - Sqr(cd.z));
        const float minS = std::min({s0, s1, s2, se01, se12, se20, sc});
        const float maxS = std::max({s0, s1, s2, se01, se12, se20, sc});
        const float poleSin = std::sin(0.5f * kPiDiv2F / img->h);
        float minSpd, maxSpd;
        SamplesPerDim(v0, v1, v2, cen, std::max(minS, poleSin), std::max(maxS, poleSin), minSpd, maxSpd);
        const uint32_t maxSpdI = (uint32_t)std::ceil(maxSpd);
        if (maxSamplesPerDim != 0 && maxSamplesPerDim < maxSpdI) return true;
        const float span = maxSpd / minSpd;
        if (span >= maxSpan && maxSpd > 32.0f) {
            if (NeedsSubdivImpl(v0, s0, e01, se01, e20, se20, whole)) return true;
            if (NeedsSubdivImpl(v1, s1, e12, se12, e01, se01, whole)) return true;
            if (NeedsSubdivImpl(v2, s2, e20, se20, e12, se12, whole)) return true;
            if (NeedsSubdivImpl(e01, se01, e12, se12, e20, se20, whole)) return true;
            return false;
        }
        return ErrorTooLarge(whole, v0, v1, v2, maxSpdI);
    }
    bool NeedsSubdiv(const Tri &t) const {            // TriangleHasToBeSubdivided, :3399-3433
        if (t.level < minLevel) return true;
        if (t.level >= maxLevel) return false;
        const V3 &d0 = verts[t.vi[0]].dir, &d1 = verts[t.vi[1]].dir, &d2 = verts[t.vi[2]].dir;
        return NeedsSubdivImpl(d0, std::sqrt(1.f - Sqr(d0.z)), d1, std::sqrt(1.f - Sqr(d1.z)), d2,
                               std::sqrt(1.f - Sqr(d2.z)), t);
    }

    void Build(const EmImage &im, float aMaxErr, uint32_t aMin, uint32_t aMax) {
        img = &im;
        maxErr = aMaxErr > 0.f ? aMaxErr : 5.f;
        minLevel = aMin;
        maxLevel = aMax;
        verts.clear();
        // GenerateInitialEmTriangulation, :2993-3025
        V3 iv[12]; uint32_t f[20][3];
        UnitIcosahedron(iv, f);
        uint32_t idx[12];
        for (int i = 0; i < 12; ++i) idx[i] = NewVertex(iv[i]);
        std::deque<Tri> todo;
        for (int i = 0; i < 20; ++i) todo.push_back(Tri{{idx[f[i][0]], idx[f[i][1]], idx[f[i][2]]}, 0});
        // RefineEmTriangulation, :3090-3155
        std::list<Tri> refined;
        while (!todo.empty()) {
            Tri t = todo.front();
            todo.pop_front();
            if (NeedsSubdiv(t)) {
                // SubdivideTriangle, :3436-3472 (new midpoint vertices per triangle)
                const V3 d0 = verts[t.vi[0]].dir, d1 = verts[t.vi[1]].dir, d2 = verts[t.vi[2]].dir;
                const uint32_t n0 = NewVertex(Normalize((d0 + d1) / 2.f));
                const uint32_t n1 = NewVertex(Normalize((d1 + d2) / 2.f));
                const uint32_t n2 = NewVertex(Normalize((d2 + d0) / 2.f));
                const uint32_t L = t.level + 1;
                const Tri sub[4] = {{{n0, n1, n2}, L}, {{t.vi[0], n0, n2}, L},
                                    {{n0, t.vi[1], n1}, L}, {{n1, t.vi[2], n2}, L}};
                for (int i = 0; i < 4; ++i) todo.push_front(sub[i]);
            } else {
                refined.push_front(t);
            }
        }
        // BuildTriangleTree, :3755-3793
        std::list<std::unique_ptr<Node>> layer;
        for (const Tri &t : refined) {
            std::unique_ptr<Node> n(new Node);
            n->leaf = true;
            n->tri = t;
            layer.push_back(std::move(n));
        }
        while (layer.size() > 1) {
            std::list<std::unique_ptr<Node>> next;
            while (!layer.empty()) {
                std::unique_ptr<Node> s(new Node);
                while (!layer.empty() && s->child.size() < 6) {
                    std::unique_ptr<Node> c = std::move(layer.front());
                    layer.pop_front();
                    s->childW.push_back(c->leaf ? TriWeight(c->tri) : NodeWeight(*c));
                    s->child.push_back(std::move(c));
                }
                next.push_back(std::move(s));
            }
            layer.swap(next);
        }
        root = std::move(layer.front());
    }

    float WholeIntegral(const SH9 &coeffs) const {     // ComputeWholeIntegral, :2416-2432
        if (!root) return 0.f;
        if (root->leaf) return Dot9(TriWeight(root->tri), coeffs);
        return Dot9(NodeWeight(*root), coeffs);
    }
    float BasePdf(const V3 &dir, const SH9 &coeffs) const {   // GetBasePdf, :2387-2413
        const float lum = img->Evaluate(dir).Luminance();
        const float cc = Dot9(ShBasis(dir), coeffs);
        const float whole = WholeIntegral(coeffs);
        if (IsTiny(whole)) return 0.f;
        return (lum * cc) / whole;
    }
    // PickTriangle, :3798-3875
    const Node *PickTriangle(const SH9 &coeffs, float &u) const {
        const Node *node = root.get();
        while (node && !node->leaf) {
            const size_t cc = node->child.size();
            float ci[6], cs[7];
            for (size_t i = 0; i < cc; ++i) ci[i] = Dot9(node->childW[i], coeffs);
            cs[0] = 0.f;
            for (size_t i = 0; i < cc; ++i) cs[i + 1] = cs[i] + ci[i];
            const float whole = cs[cc];
            if (whole > 1E-10f) {
                const float inv = 1.f / whole;
                for (size_t i = 1; i <= cc; ++i) cs[i] *= inv;
            } else {
                for (size_t i = 0; i < cc; ++i) ci[i] = 1.f / cc;   // reference writes the unused array
            }
            const float ut = u * 0.999999f;
            float *end = cs + cc + 1;
            float *ub = std::upper_bound(cs, end, ut);
            if (ub == end) --ub;
            const size_t child = (size_t)(ub - cs) - 1u;
            const float low = *(ub - 1);
            const float integ = *ub - low;
            const float nu = (ut - low) / std::max(integ, 0.00001f);
            u = std::min(nu, 1.f);
            node = node->child[child].get();
        }
        return node;
    }
    // SampleTriangleFFunction / GFunction / Bilinear, :3709-3747, :3882-3905; math.hxx:284-342
    static float NewtonCubic(float a, float b, float c, float d) {
        float x = 0.5f, fval, der;
        for (int it = 0; it < 4; ++it) {
            fval = x * (x * (x * a + b) + c) + d;
            der = x * (x * 3.f * a + 2.f * b) + c;
            if (std::abs(der) < 0.0001f) return 0.5f;
            x = x - fval / der;
            x = Clamp(x, 0.f, 1.f);
        }
        return x;
    }
    // s (SampleTriangleFFunction) and t (SampleTriangleGFunction) of the barycentric sample
    static void SampleTriST(float ux, float uy, float va, float vb, float vc, float &s, float &t) {
        const float x = (vb - va) / 3 + (vc - vb) / 6;
        const float y = va / 2;
        const float alpha = x / (x + y), beta = y / (x + y);
        s = NewtonCubic(alpha, beta, 0, -ux);
        const float t0 = (s * (vc - vb)) + (2 * (1 - s) * va) + (s * vb);
        const float gamma = s * (vc - vb) / t0;
        const float rho = 2 * ((1 - s) * va + s * vb) / t0;
        const float discr = rho * rho + 4 * gamma * uy;
        t = (2 * uy) / (rho + SafeSqrt(discr));
    }
    static void SampleTriBilinear(float ux, float uy, float va, float vb, float vc, float &b0, float &b1) {
        float s, t;
        SampleTriST(ux, uy, va, vb, vc, s, t);
        b0 = float(1 - s);
        b1 = float(s * (1 - t));
    }
    // SampleHemisphereImpl, :2162-2217
    bool SampleHemisphere(const Frame &fr, const SH9 &coeffs, float u0, float u1, V3 &dir, float &pdf,
                          V3 &radCos) const {
        if (!root) return false;
        float ux = u0;
        const Node *tn = PickTriangle(coeffs, ux);
        if (!tn) return false;
        const Tri &t = tn->tri;
        const float v0 = Dot9(verts[t.vi[0]].w, coeffs), v1 = Dot9(verts[t.vi[1]].w, coeffs),
                    v2 = Dot9(verts[t.vi[2]].w, coeffs);
        float b0, b1;
        SampleTriBilinear(ux, u1, v0, v1, v2, b0, b1);
        const float b2 = (1.0f - b0 - b1);
        const float val = b0 * v0 + b1 * v1 + b2 * v2;
        V3 p = verts[t.vi[0]].dir * b0 + verts[t.vi[1]].dir * b1 + verts[t.vi[2]].dir * b2;
        dir = p / p.Length();
        const float whole = WholeIntegral(coeffs);
        pdf = IsTiny(whole) ? 0.f : val / whole;
        const V3 flipped = -dir;
        pdf += BasePdf(flipped, coeffs);
        float c = Dot(dir, fr.Normal());
        if (c < 0.f) { dir = flipped; c *= -1.f; }
        radCos = img->Evaluate(dir) * c;
        return true;
    }
    // SampleImpl, :2084-2157 (uSide used only when both sides are sampled)
    bool Sample(const Frame &fr, bool front, bool back, float uSide, float u0, float u1, V3 &dir,
                float &pdf, V3 &radCos) const {
        if (front && back) {
            Frame lf = fr; lf.SwitchNormal();
            const SH9 uc = ClampedCosCoeffs(fr.Normal(), true), lc = ClampedCosCoeffs(lf.Normal(), true);
            const float ui = WholeIntegral(uc), li = WholeIntegral(lc), wi = ui + li;
            if (IsTiny(wi)) return false;
            if (uSide * wi < ui) {
                if (!SampleHemisphere(fr, uc, u0, u1, dir, pdf, radCos)) return false;
                pdf *= ui / wi;
            } else {
                if (!SampleHemisphere(lf, lc, u0, u1, dir, pdf, radCos)) return false;
                pdf *= li / wi;
            }
            return true;
        } else if (front) {
            return SampleHemisphere(fr, ClampedCosCoeffs(fr.Normal(), true), u0, u1, dir, pdf, radCos);
        } else if (back) {
            Frame lf = fr; lf.SwitchNormal();
            return SampleHemisphere(lf, ClampedCosCoeffs(lf.Normal(), true), u0, u1, dir, pdf, radCos);
        }
        return false;
    }
    float PdfHemisphere(const V3 &dir, const Frame &fr, const SH9 &coeffs) const {   // :2281-2299
        if (Dot(dir, fr.Normal()) < 0.f) return 0.f;
        return BasePdf(dir, coeffs) + BasePdf(-dir, coeffs);
    }
    float Pdf(const V3 &dir, const Frame &fr, bool front, bool back) const {          // :2220-2278
        if (front && back) {
            Frame lf = fr; lf.SwitchNormal();
            const SH9 uc = ClampedCosCoeffs(fr.Normal(), true), lc = ClampedCosCoeffs(lf.Normal(), true);
            const float ui = WholeIntegral(uc), li = WholeIntegral(lc), wi = ui + li;
            if (IsTiny(wi)) return 0.f;
            if (Dot(dir, fr.Normal()) >= 0.f) return PdfHemisphere(dir, fr, uc) * (ui / wi);
            return PdfHemisphere(dir, lf, lc) * (li / wi);
        } else if (front) {
            return PdfHemisphere(dir, fr, ClampedCosCoeffs(fr.Normal(), true));
        } else if (back) {
            Frame lf = fr; lf.SwitchNormal();
            return PdfHemisphere(dir, lf, ClampedCosCoeffs(lf.Normal(), true));
        }
        return 0.f;
    }
    float EstimateIrradiance(const Frame &fr, bool front, bool back) const {          // :2304-2358
        Frame lf = fr; lf.SwitchNormal();
        const float ui = WholeIntegral(ClampedCosCoeffs(fr.Normal(), true));
        const float li = WholeIntegral(ClampedCosCoeffs(lf.Normal(), true));
        if (front && back) return ui + li;
        if (front) return ui;
        if (back) return li;
        return 0.f;
    }
    void Leaves(std::vector<const Node *> &out) const {
        // pre-order traversal in child order
        std::function<void(const Node *)> rec = [&](const Node *n) {
            if (n->leaf) { out.push_back(n); return; }
            for (auto &c : n->child) rec(c.get());
        };
        if (root) rec(root.get());
    }
};

}  // namespace orc
