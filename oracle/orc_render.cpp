// Oracle (test infrastructure, not the product): lights, scene queries, the reference renderers
// (path tracer with NEE+MIS, naive path tracer, direct illumination) and the render loop,
// restated on the CPU.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// this library.
#include <omp.h>

#include <chrono>
#include <cstdio>
#include <string>

#include "orc_em.hpp"

namespace orc {

struct LightSample {      // light_sample.hxx:11-31
    V3 sample;
    float pdfW = 0.f, lightProb = 1.f;
    V3 wig;
    float dist = 0.f;
    bool IsPoint() const { return pdfW == InfF(); }
};

struct Light {
    int type = PG3G_LIGHT_AREA;
    // AreaLight, lights.hxx:55-229
    V3 p0, e1, e2;
    Frame frame;
    V3 radiance;           // area: Le; point: intensity; background: constant radiance
    float area = 0.f, invArea = 0.f;

    void InitArea(const V3 &a, const V3 &b, const V3 &c, const V3 &power) {   // :62-76, :104-112
        p0 = a; e1 = b - a; e2 = c - a;
        V3 n = Cross(e1, e2);
        const float len = n.Length();
        area = len / 2.f;
        invArea = 2.f / len;
        frame.SetFromZ(n / n.Length());
        radiance = power * (invArea / kPiF);
    }
};

struct Sphere { V3 c; float r; int mat; };
struct Tri { V3 p[3]; V3 n; int mat; };
struct Isect { float dist; int matID, lightID; V3 normal; };

struct Scene {
    std::vector<Tri> tris;
    std::vector<Sphere> sphs;
    std::vector<Material> mats;
    std::vector<Light> lights;
    int bgLight = -1;
    EmImage em;
    int emSampler = PG3G_EM_NONE;
    SimpleSampler simple;
    SteerableSampler steer;
    // camera (camera.hxx:13-100)
    V3 camPos;
    double r2w[16];   // raster->world, row-major
    uint32_t resX = 0, resY = 0;
    pg3g_render_params params{};

    // ---------------------------------------------------------------- geometry (scene_graph.hxx)
    static bool TriIntersect(const Tri &t, const V3 &org, const V3 &dir, float tmin, Isect &r) {  // :103-134
        const V3 ao = t.p[0] - org, bo = t.p[1] - org, co = t.p[2] - org;
        const V3 v0 = Cross(co, bo), v1 = Cross(bo, ao), v2 = Cross(ao, co);
        const float v0d = Dot(v0, dir), v1d = Dot(v1, dir), v2d = Dot(v2, dir);
        if (((v0d < 0.f) && (v1d < 0.f) && (v2d < 0.f)) || ((v0d >= 0.f) && (v1d >= 0.f) && (v2d >= 0.f))) {
            const float d = Dot(t.n, ao) / Dot(t.n, dir);
            if ((d > tmin) & (d < r.dist)) { r.normal = t.n; r.matID = t.mat; r.dist = d; return true; }
        }
        return false;
    }
    static bool SphIntersect(const Sphere &s, const V3 &org, const V3 &dir, float tmin, Isect &r) {  // :176-245
        const V3 o = org - s.c;
        const float A = Dot(dir, dir);
        const float B = 2 * Dot(dir, o);
        const float C = Dot(o, o) - (s.r * s.r);
        // B*B - 4*A*C has only float operands: it is evaluated in FP32 and then widened; the
        // square root, q and both roots are FP64 (scene_graph.hxx:193-215).
        const double disc = B * B - 4 * A * C;
        if (disc < 0) return false;
        const double ds = std::sqrt(disc);
        const double q = (B >= 0) ? ((-B - ds) / 2.f) : ((-B + ds) / 2.f);
        double t0 = q / A;
        double t1 = (std::abs(q) > 0.1) ? (C / q) : t0;
        if (t0 > t1) std::swap(t0, t1);
        float resT;
        if (t0 > tmin && t0 < r.dist) resT = float(t0);
        else if (t1 > tmin && t1 < r.dist) resT = float(t1);
        else return false;
        r.dist = resT;
        r.matID = s.mat;
        r.normal = Normalize(o + V3(resT) * dir);
        return true;
    }
    bool Intersect(const V3 &org, const V3 &dir, float tmin, Isect &r) const {   // scene.hxx:42-59
        bool any = false;
        for (const Tri &t : tris) any |= TriIntersect(t, org, dir, tmin, r);
        for (const Sphere &s : sphs) any |= SphIntersect(s, org, dir, tmin, r);
        if (any) r.lightID = mats[r.matID].lightId;
        return any;
    }
    bool Occluded(const V3 &p, const V3 &dir, float tmax) const {               // scene.hxx:61-74
        const V3 org = p + dir * kEpsRay;
        Isect r;
        r.dist = tmax - 2 * kEpsRay;
        for (const Tri &t : tris) if (TriIntersect(t, org, dir, 0.f, r)) return true;
        for (const Sphere &s : sphs) if (SphIntersect(s, org, dir, 0.f, r)) return true;
        return false;
    }

    // ---------------------------------------------------------------- camera
    void GenerateRay(float x, float y, V3 &org, V3 &dir) const {                // camera.hxx:78-100
        const double v[3] = {x, y, 0.0};
        float w = (float)r2w[15];
        for (int c = 0; c < 3; ++c) w += (float)r2w[12 + c] * (float)v[c];
        const float invW = 1.f / w;
        V3 p;
        for (int rr = 0; rr < 3; ++rr) {
            float a = (float)r2w[rr * 4 + 3];
            for (int c = 0; c < 3; ++c) a += (float)v[c] * (float)r2w[rr * 4 + c];
            p[rr] = a * invW;
        }
        org = camPos;
        dir = Normalize(p - camPos);
    }

    // ---------------------------------------------------------------- lights (lights.hxx)
    void AreaComputeSample(const Light &L, const V3 &surfPt, const V3 &samplePt, const Frame &sf,
                           const Material &m, LightSample &o) const {           // :171-219
        o.wig = samplePt - surfPt;
        const float d2 = o.wig.LenSqr();
        o.dist = std::sqrt(d2);
        o.wig = o.wig / o.dist;
        const float cosOut = -Dot(L.frame.Normal(), o.wig);
        float cosIn = Dot(sf.Normal(), o.wig);
        if (m.front && m.back) cosIn = std::abs(cosIn);
        else if (m.front) cosIn = std::max(cosIn, 0.0f);
        else if (m.back) cosIn = std::max(-cosIn, 0.0f);
        else cosIn = 0.0f;
        o.sample = (cosOut > 0.f) ? L.radiance * cosIn : V3(0.f);
        o.pdfW = std::max(L.invArea * (d2 / std::abs(cosOut)), kEpsDist);
        o.lightProb = 1.0f;
    }
    void PointCompute(const Light &L, const V3 &surfPt, const Frame &sf, const Material &m,
                      LightSample &o) const {                                   // :296-331
        o.wig = L.p0 - surfPt;
        const float d2 = o.wig.LenSqr();
        o.dist = std::sqrt(d2);
        o.wig = o.wig / o.dist;
        float cosIn = Dot(sf.Normal(), o.wig);
        if (m.front && m.back) cosIn = std::abs(cosIn);
        else if (m.front) cosIn = std::max(cosIn, 0.0f);
        else if (m.back) cosIn = std::max(-cosIn, 0.0f);
        else cosIn = 0.0f;
        o.sample = (cosIn > 0.f) ? L.radiance * cosIn / d2 : V3(0.f);
        o.pdfW = InfF();
        o.lightProb = 1.0f;
    }
    // EnvironmentMap::PdfW via the active sampler (em.hxx:113-120)
    float EmPdf(const V3 &dir, const Frame &sf, bool front, bool back) const {
        switch (emSampler) {
        case PG3G_EM_SIMPLE: return simple.Pdf(dir);
        case PG3G_EM_STEERABLE: return steer.Pdf(dir, sf, front, back);
        default: return CosSpherePdfW(front, back, sf.ToLocal(dir));
        }
    }
    // BackgroundLight::GetEmmision(dir, pdf, frame, material), lights.hxx:379-407
    V3 BgEmission(const V3 &wig, float *pdf, const Frame *sf, const Material *m) const {
        const Light &L = lights[bgLight];
        const bool front = m ? m->front : false, back = m ? m->back : false;
        if (em.w > 0) {
            const V3 rad = em.Evaluate(wig);
            if (pdf && sf) *pdf = EmPdf(wig, *sf, front, back);
            return rad;
        }
        if (pdf && sf && m) *pdf = CosSpherePdfW(front, back, sf->ToLocal(wig));
        return L.radiance;
    }
    // AbstractLight::GetEmmision for a hit on light geometry
    V3 Emission(int lightId, const V3 &lightPt, const V3 &wol, const V3 &surfPt, float *pdf,
                const Frame *sf, const Material *m) const {
        const Light &L = lights[lightId];
        if (L.type == PG3G_LIGHT_AREA) {                                        // :78-102
            if (pdf) {
                V3 wig = lightPt - surfPt;
                const float d2 = wig.LenSqr();
                wig = wig / std::sqrt(d2);
                const float ac = std::abs(Dot(L.frame.Normal(), wig));
                *pdf = std::max(L.invArea * (d2 / ac), kEpsDist);
            }
            return (wol.z <= 0.) ? V3(0.f) : L.radiance;
        }
        if (L.type == PG3G_LIGHT_POINT) { if (pdf) *pdf = InfF(); return V3(0.f); }
        return BgEmission(-wol, pdf, sf, m);                                    // :410-420
    }
    float EmEstimateIrradiance(const Frame &sf, bool front, bool back, const PathRng &rng,
                               uint32_t vtx) const {                            // em.hxx:123-170
        if (emSampler == PG3G_EM_STEERABLE) return steer.EstimateIrradiance(sf, front, back);
        float sum = 0.f;
        for (uint32_t round = 0; round < 10; round++) {
            const uint32_t d = kDimEstimateBase + round * 8;
            float p1c;
            const V3 w1 = SampleCosSphereParamPdfW(rng.Get(vtx, 0, d), rng.Get(vtx, 0, d + 1),
                                                   rng.Get(vtx, 0, d + 2), front, back, p1c);
            const V3 g1 = sf.ToWorld(w1);
            const V3 s1 = em.Evaluate(g1) * std::abs(w1.z);
            const float p1s = simple.Pdf(g1);
            V3 g2, s2; float p2s;
            simple.Sample(sf, front, back, rng.Get(vtx, 0, d + 3), rng.Get(vtx, 0, d + 4), g2, p2s, s2);
            const float p2c = CosSpherePdfW(front, back, sf.ToLocal(g2));
            sum += s1.Luminance() / (p1c + p1s) + s2.Luminance() / (p2c + p2s);
        }
        return sum / 10;
    }
    float EstimateContribution(int li, const V3 &surfPt, const Frame &sf, const Material &m,
                               const PathRng &rng, uint32_t vtx) const {
        const Light &L = lights[li];
        if (L.type == PG3G_LIGHT_AREA) {                                        // :133-167
            const V3 P1 = L.p0 + L.e1, P2 = L.p0 + L.e2;
            const V3 P3 = L.p0 + 0.33f * L.e1 + 0.33f * L.e2;
            LightSample s0, s1, s2, s3;
            AreaComputeSample(L, surfPt, L.p0, sf, m, s0);
            AreaComputeSample(L, surfPt, P1, sf, m, s1);
            AreaComputeSample(L, surfPt, P2, sf, m, s2);
            AreaComputeSample(L, surfPt, P3, sf, m, s3);
            return (s0.sample.Luminance() / s0.pdfW + s1.sample.Luminance() / s1.pdfW +
                    s2.sample.Luminance() / s2.pdfW + s3.sample.Luminance() / s3.pdfW) / 4.f;
        }
        if (L.type == PG3G_LIGHT_POINT) {                                      // :277-290
            LightSample s; PointCompute(L, surfPt, sf, m, s);
            return s.sample.Luminance();
        }
        if (em.w > 0) return EmEstimateIrradiance(sf, m.front, m.back, rng, vtx);   // :451-467
        return L.radiance.Luminance() * kPiF;
    }
    bool SampleIllumination(int li, const V3 &surfPt, const Frame &sf, const Material &m,
                            const PathRng &rng, uint32_t vtx, uint32_t sub, LightSample &o) const {
        const Light &L = lights[li];
        const float u0 = rng.Get(vtx, sub, kDimLight0), u1 = rng.Get(vtx, sub, kDimLight1);
        if (L.type == PG3G_LIGHT_AREA) {                                        // :114-128
            const V3 sp = SampleUniformTriangle(L.p0, L.p0 + L.e1, L.p0 + L.e2, u0, u1);
            AreaComputeSample(L, surfPt, sp, sf, m, o);
            return true;
        }
        if (L.type == PG3G_LIGHT_POINT) { PointCompute(L, surfPt, sf, m, o); return true; }
        const float u2 = rng.Get(vtx, sub, kDimLight2);
        o.lightProb = 1.0f;                                                     // em_sampler.hxx:49-64
        o.dist = std::numeric_limits<float>::max();
        if (em.w > 0) {
            switch (emSampler) {
            case PG3G_EM_SIMPLE:
                return simple.Sample(sf, m.front, m.back, u0, u1, o.wig, o.pdfW, o.sample);
            case PG3G_EM_STEERABLE:
                return steer.Sample(sf, m.front, m.back, u2, u0, u1, o.wig, o.pdfW, o.sample);
            default: {
                const V3 w = SampleCosSphereParamPdfW(u0, u1, u2, m.front, m.back, o.pdfW);
                o.wig = sf.ToWorld(w);
                o.sample = em.Evaluate(o.wig) * std::abs(w.z);
                return true;
            }
            }
        }
        // constant background: cosine sampler (em_cosine_sampler.hxx:12-33)
        const V3 w = SampleCosSphereParamPdfW(u0, u1, u2, m.front, m.back, o.pdfW);
        o.wig = sf.ToWorld(w);
        o.sample = L.radiance * std::abs(w.z);
        return true;
    }
};

// -------------------------------------------------------------------------------------------------
// Renderers (path_tracer_base.hxx, path_tracer.hxx, direct_illumination.hxx)
struct Renderer {
    const Scene &S;
    explicit Renderer(const Scene &s) : S(s) {}

    struct LightCtx { std::vector<float> cache; bool valid = false; };      // path_tracer_base.hxx:22-32

    static float MISWeight2(float p1, uint32_t n1, float p2, uint32_t n2) {  // :522-541 (power)
        const float s1 = n1 * p1, s2 = n2 * p2;
        const float q1 = s1 * s1, q2 = s2 * s2;
        return q1 / (q1 + q2) / n1;
    }
    void FillCache(const V3 &p, const Frame &f, const Material &m, LightCtx &ctx, const PathRng &rng,
                   uint32_t vtx) const {
        if (ctx.valid) return;
        ctx.cache.resize(S.lights.size());
        for (size_t i = 0; i < S.lights.size(); ++i) ctx.cache[i] = S.EstimateContribution((int)i, p, f, m, rng, vtx);
        ctx.valid = true;
    }
    void PickSingleLight(const V3 &p, const Frame &f, const Material &m, LightCtx &ctx, const PathRng &rng,
                         uint32_t vtx, uint32_t sub, int &id, float &prob) const {   // :191-257
        const size_t n = S.lights.size();
        if (n == 0) { id = -1; return; }
        if (n == 1) { id = 0; prob = 1.f; return; }
        FillCache(p, f, m, ctx, rng, vtx);
        std::vector<float> cdf(n + 1);
        float sum = 0.f;
        cdf[0] = 0.f;
        for (size_t i = 0; i < n; i++) { sum += ctx.cache[i]; cdf[i + 1] = sum; }
        const float u = rng.Get(vtx, sub, kDimLightPick);
        if (sum > 0.f) {
            const float rv = u * sum;
            uint32_t li = 0;
            for (; (rv > cdf[li + 1]) && (li < n); li++) {}
            id = (int)li;
            prob = (cdf[li + 1] - cdf[li]) / sum;
        } else {
            id = std::min((int)(u * n), (int)(n - 1));
            prob = 1.f / n;
        }
    }
    float LightPickingProbability(const V3 &p, const Frame &f, const Material &m, int lightId,
                                  LightCtx &ctx, const PathRng &rng, uint32_t vtx) const {  // :260-313
        const size_t n = S.lights.size();
        if (n == 0) return 0.f;
        if (n == 1) return 1.f;
        FillCache(p, f, m, ctx, rng, vtx);
        float sum = 0.f, est = 0.f;
        for (size_t i = 0; i < n; i++) { if ((int)i == lightId) est = ctx.cache[i]; sum += ctx.cache[i]; }
        return (sum > 0.f) ? est / sum : 1.f / n;
    }
    bool SampleLightsSingle(const V3 &p, const Frame &f, const Material &m, LightCtx &ctx,
                            const PathRng &rng, uint32_t vtx, uint32_t sub, LightSample &o) const {  // :156-188
        int id = -1; float prob = 0.f;
        PickSingleLight(p, f, m, ctx, rng, vtx, sub, id, prob);
        if (id < 0) return false;
        const bool ok = S.SampleIllumination(id, p, f, m, rng, vtx, sub, o);
        o.lightProb = prob;
        return ok;
    }
    void AddMISLightSample(const LightSample &ls, uint32_t nL, uint32_t nB, const V3 &p, const Frame &f,
                           const V3 &wol, const Material &m, V3 &out) const {    // :350-425
        if (ls.sample.Max() <= 0.) return;
        if (S.Occluded(p, ls.wig, ls.dist)) return;
        MatRec r; r.wil = f.ToLocal(ls.wig); r.wol = wol;
        m.Eval(r);
        if (!ls.IsPoint()) {
            const float bsdfPdf = r.pdfW * r.compProb;
            const float lightPdf = ls.pdfW * ls.lightProb;
            out += (ls.sample * r.att * MISWeight2(lightPdf, nL, bsdfPdf, nB)) / lightPdf;
        } else {
            out += (ls.sample * r.att) / (ls.lightProb * nL);
        }
    }
    void AddSingleLightSample(const LightSample &ls, const V3 &p, const Frame &f, const Material &m,
                              const V3 &wol, V3 &out) const {                     // :315-348
        if (ls.sample.Max() <= 0.) return;
        if (S.Occluded(p, ls.wig, ls.dist)) return;
        MatRec r; r.wil = f.ToLocal(ls.wig); r.wol = wol;
        m.Eval(r);
        if (!ls.IsPoint()) out += ls.sample * r.att / (ls.pdfW * ls.lightProb);
        else out += ls.sample * r.att / ls.lightProb;
    }
    // GetDirectRadianceFromDirection, :92-154
    V3 DirectFromDir(const V3 &p, const Frame &f, const Material &m, const V3 &wil, LightCtx &ctx,
                     const PathRng &rng, uint32_t vtx, float *pdf, float *pick) const {
        int lightId = -1;
        V3 Li(0.f);
        const V3 wig = f.ToWorld(wil);
        const float tmin = EpsRayCos(wil.z);
        Isect is; is.dist = 1e36f;
        if (S.Intersect(p, wig, tmin, is)) {
            if (is.lightID >= 0) {
                const V3 lp = p + wig * is.dist;
                Frame lf; lf.SetFromZ(is.normal);
                Li = S.Emission(is.lightID, lp, lf.ToLocal(-wig), p, pdf, &f, &m);
                lightId = is.lightID;
            }
        } else if (S.bgLight >= 0) {
            Li = S.BgEmission(wig, pdf, &f, &m);
            lightId = S.bgLight;
        }
        if (pick && lightId >= 0) *pick = LightPickingProbability(p, f, m, lightId, ctx, rng, vtx);
        return Li;
    }

    // PathTracer::EstimateIncomingRadiancePT, path_tracer.hxx:167-430
    void PT(const V3 &org, const V3 &dir, float tmin, uint32_t len, bool computeRefl, float budget,
            const PathRng &rng, V3 &emitted, V3 &reflected, const Frame *prevF, const Material *prevM,
            float *lightPdf, int *lightIdOut) const {
        const pg3g_render_params &P = S.params;
        emitted = V3(0.f);
        reflected = V3(0.f);
        LightCtx ctx;
        Isect is; is.dist = 1e36f;
        if (S.Intersect(org, dir, tmin, is)) {
            const V3 sp = org + dir * is.dist;
            Frame f; f.SetFromZ(is.normal);
            const V3 wol = f.ToLocal(-dir);
            const Material &m = S.mats[is.matID];
            if (is.lightID >= 0 && len >= P.min_path_length) {
                emitted += S.Emission(is.lightID, sp, wol, org, lightPdf, prevF, prevM);
                if (lightIdOut) *lightIdOut = is.lightID;
            }
            if (!computeRefl) return;
            if (m.ReflectanceZero()) return;
            if (P.max_path_length > 0 && len >= P.max_path_length) return;
            if (len >= 1200) return;
            // ComputeSplittingCounts, :432-454
            const float lvl = P.splitting_level;
            const uint32_t nB = (uint32_t)std::lroundf((1.f - lvl) * 1.f + lvl * budget);
            const uint32_t nL = (uint32_t)std::lroundf(std::max(P.light_to_bsdf_ratio * nB, 1.f));
            const float nextBudget = 1 + ((budget - nB) / (float)nB);
            if ((len + 1) >= P.min_path_length) {
                for (uint32_t s = 0; s < nL; s++) {
                    LightSample ls;
                    if (SampleLightsSingle(sp, f, m, ctx, rng, len, s, ls))
                        AddMISLightSample(ls, nL, nB, sp, f, wol, m, reflected);
                }
            }
            float rr = 1.0f;
            if (P.max_path_length == 0) rr = Clamp(m.RRProb(wol), 0.0f, 0.997f);
            for (uint32_t s = 0; s < nB; s++) {
                bool cut = false;
                if (P.max_path_length == 0) {
                    if (rng.Get(len, s, kDimRR) > rr) cut = true;
                }
                MatRec r; r.wol = wol;
                m.Sample(rng.Get(len, s, kDimBsdf0), rng.Get(len, s, kDimBsdf1), rng.Get(len, s, kDimBsdf2), r);
                if (r.IsBlocker()) continue;
                V3 be, br;
                float blPdf = 0.f;
                int blId = -1;
                const V3 wig = f.ToWorld(r.wil);
                PathRng child = rng;
                if (nB > 1) child.branch = (rng.branch * 7u + s + 1u) & 0xFFFFu;
                PT(sp, wig, EpsRayCos(r.CosAbs()), len + 1, !cut, nextBudget, child, be, br, &f, &m,
                   &blPdf, &blId);
                if (!be.IsZero() && blId >= 0) {
                    if (r.IsFinite()) {
                        const float pick = LightPickingProbability(sp, f, m, blId, ctx, rng, len);
                        const float lp = blPdf * pick;
                        const float bp = r.pdfW * r.compProb;
                        reflected += (r.att * r.CosAbs() * be * MISWeight2(bp, nB, lp, nL)) / bp;
                    } else {
                        reflected += (r.att * be) / (static_cast<float>(nB) * r.compProb);
                    }
                }
                if (!cut && !br.IsZero()) {
                    V3 ind;
                    if (r.IsFinite()) ind = (r.att * r.CosAbs() * br) / (r.pdfW * nB * rr * r.compProb);
                    else ind = (r.att * br) / (nB * rr * r.compProb);
                    if (P.indirect_clipping > 0.f) ind.ClipProportionally(P.indirect_clipping);
                    reflected += ind;
                }
            }
        } else {
            if (len >= P.min_path_length && S.bgLight >= 0) {
                emitted += S.BgEmission(dir, lightPdf, prevF, prevM);
                if (lightIdOut) *lightIdOut = S.bgLight;
            }
        }
    }

    // PathTracer::EstimateIncomingRadiancePTNaive, path_tracer.hxx:45-165
    V3 PTNaive(V3 org, V3 dir, float tmin, const PathRng &rng) const {
        const pg3g_render_params &P = S.params;
        V3 out(0.f), thr(1.0f);
        for (uint32_t len = 1;;) {
            Isect is; is.dist = 1e36f;
            if (!S.Intersect(org, dir, tmin, is)) {
                if (len >= P.min_path_length && S.bgLight >= 0) out += S.BgEmission(dir, nullptr, nullptr, nullptr) * thr;
                return out;
            }
            const V3 sp = org + dir * is.dist;
            Frame f; f.SetFromZ(is.normal);
            const V3 wol = f.ToLocal(-dir);
            const Material &m = S.mats[is.matID];
            if (is.lightID >= 0 && len >= P.min_path_length)
                out += S.Emission(is.lightID, sp, wol, V3(0.f), nullptr, nullptr, nullptr) * thr;
            if (m.ReflectanceZero()) return out;
            if (P.max_path_length > 0 && len >= P.max_path_length) return out;
            if (len >= 1200) return out;
            float rr = 1.0f;
            if (P.max_path_length == 0) {
                rr = Clamp(m.RRProb(wol), 0.0f, 0.997f);
                if (rng.Get(len, 0, kDimRR) > rr) return out;
            }
            MatRec r; r.wol = wol;
            m.Sample(rng.Get(len, 0, kDimBsdf0), rng.Get(len, 0, kDimBsdf1), rng.Get(len, 0, kDimBsdf2), r);
            if (r.IsBlocker()) return out;
            org = sp;
            dir = f.ToWorld(r.wil);
            tmin = EpsRayCos(r.CosAbs());
            const V3 seg = r.IsFinite() ? (r.att * r.CosAbs()) / (r.pdfW * rr * r.compProb)
                                        : r.att / (rr * r.compProb);
            thr *= seg;
            len++;
        }
    }

    // DirectIllumination::EstimateIncomingRadiance, direct_illumination.hxx:15-156
    V3 Direct(int alg, const V3 &org, const V3 &dir, const PathRng &rng) const {
        LightCtx ctx;
        Isect is; is.dist = 1e36f;
        if (!S.Intersect(org, dir, 0.f, is)) {
            return (S.bgLight >= 0) ? S.BgEmission(dir, nullptr, nullptr, nullptr) : V3(0.f);
        }
        const V3 sp = org + dir * is.dist;
        Frame f; f.SetFromZ(is.normal);
        const V3 wol = f.ToLocal(-dir);
        const Material &m = S.mats[is.matID];
        const uint32_t vtx = 1;
        V3 Lo(0.f);
        if (alg == PG3G_ALG_DIRECT_LIGHT_ALL) {
            for (size_t i = 0; i < S.lights.size(); i++) {
                LightSample ls;
                if (S.SampleIllumination((int)i, sp, f, m, rng, vtx, (uint32_t)i, ls))
                    AddSingleLightSample(ls, sp, f, m, wol, Lo);
            }
        } else if (alg == PG3G_ALG_DIRECT_LIGHT_SINGLE) {
            LightSample ls;
            if (SampleLightsSingle(sp, f, m, ctx, rng, vtx, 0, ls)) AddSingleLightSample(ls, sp, f, m, wol, Lo);
        } else if (alg == PG3G_ALG_DIRECT_BSDF) {
            MatRec r; r.wol = wol;
            m.Sample(rng.Get(vtx, 0, kDimBsdf0), rng.Get(vtx, 0, kDimBsdf1), rng.Get(vtx, 0, kDimBsdf2), r);
            if (!r.IsBlocker()) {
                const V3 Li = DirectFromDir(sp, f, m, r.wil, ctx, rng, vtx, nullptr, nullptr);
                Lo = r.IsFinite() ? (r.att * r.CosAbs() * Li) / (r.pdfW * r.compProb) : (r.att * Li) / r.compProb;
            }
        } else if (alg == PG3G_ALG_DIRECT_MIS) {
            LightSample ls;
            if (SampleLightsSingle(sp, f, m, ctx, rng, vtx, 0, ls)) AddMISLightSample(ls, 1, 1, sp, f, wol, m, Lo);
            MatRec r; r.wol = wol;
            m.Sample(rng.Get(vtx, 0, kDimBsdf0), rng.Get(vtx, 0, kDimBsdf1), rng.Get(vtx, 0, kDimBsdf2), r);
            if (!r.IsBlocker()) {                                               // :427-490
                float lpdf = 0.f, pick = 0.f;
                const V3 Li = DirectFromDir(sp, f, m, r.wil, ctx, rng, vtx, &lpdf, &pick);
                if (Li.Max() > 0.f) {
                    if (r.IsFinite()) {
                        const float bp = r.pdfW * r.compProb;
                        Lo += (r.att * r.CosAbs() * Li * MISWeight2(bp, 1, lpdf * pick, 1)) / bp;
                    } else {
                        Lo += (r.att * Li) / (1.f * r.compProb);
                    }
                }
            }
        }
        V3 Le(0.f);
        if (is.lightID >= 0) Le = S.Emission(is.lightID, sp, wol, V3(0.f), nullptr, nullptr, nullptr);
        return Le + Lo;
    }

    V3 Radiance(const V3 &org, const V3 &dir, const PathRng &rng) const {
        const int alg = S.params.algorithm;
        if (alg == PG3G_ALG_PATH_TRACING) {
            V3 e, r;
            PT(org, dir, 0.f, 1, true, S.params.splitting_budget, rng, e, r, nullptr, nullptr, nullptr, nullptr);
            return e + r;
        }
        if (alg == PG3G_ALG_PATH_TRACING_NAIVE) return PTNaive(org, dir, 0.f, rng);
        return Direct(alg, org, dir, rng);
    }
};

// Camera::Setup, camera.hxx:13-59 (computed in FP64)
static void Mat4Mul(const double *a, const double *b, double *o) {
    double t[16];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            double s = 0;
            for (int i = 0; i < 4; ++i) s += a[r * 4 + i] * b[i * 4 + c];
            t[r * 4 + c] = s;
        }
    std::memcpy(o, t, sizeof(t));
}
static bool Mat4Inv(const double *m, double *o) {
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 8; ++c) a[r][c] = (c < 4) ? m[r * 4 + c] : (c - 4 == r ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int r = c + 1; r < 4; ++r) if (std::abs(a[r][c]) > std::abs(a[piv][c])) piv = r;
        if (a[piv][c] == 0.0) return false;
        if (piv != c) for (int k = 0; k < 8; ++k) std::swap(a[c][k], a[piv][k]);
        const double d = a[c][c];
        for (int k = 0; k < 8; ++k) a[c][k] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            const double f = a[r][c];
            for (int k = 0; k < 8; ++k) a[r][k] -= f * a[c][k];
        }
    }
    for (int r = 0; r < 4; ++r) for (int c = 0; c < 4; ++c) o[r * 4 + c] = a[r][c + 4];
    return true;
}
static void CameraSetup(Scene &S, const pg3g_camera_desc &c) {
    const double pi = 3.14159265358979323846;
    auto nrm = [](double v[3]) { double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); for (int i = 0; i < 3; ++i) v[i] /= l; };
    auto cross = [](const double a[3], const double b[3], double o[3]) {
        o[0] = a[1] * b[2] - a[2] * b[1]; o[1] = a[2] * b[0] - a[0] * b[2]; o[2] = a[0] * b[1] - a[1] * b[0]; };
    double fwd[3] = {c.forward[0], c.forward[1], c.forward[2]}; nrm(fwd);
    const double nf[3] = {-fwd[0], -fwd[1], -fwd[2]};
    const double upIn[3] = {c.up[0], c.up[1], c.up[2]};
    double up[3]; cross(upIn, nf, up); nrm(up);
    double left[3]; cross(nf, up, left);
    const double P[3] = {c.position[0], c.position[1], c.position[2]};
    const double pos[3] = {up[0] * P[0] + up[1] * P[1] + up[2] * P[2],
                           left[0] * P[0] + left[1] * P[1] + left[2] * P[2],
                           nf[0] * P[0] + nf[1] * P[1] + nf[2] * P[2]};
    double w2c[16] = {up[0], up[1], up[2], -pos[0], left[0], left[1], left[2], -pos[1],
                      nf[0], nf[1], nf[2], -pos[2], 0, 0, 0, 1};
    const double rx = c.res_x, ry = c.res_y;
    const double half = (0.5 * c.hfov_deg) * pi / 180.0;
    const double pixelSize = std::tan(half) / (0.5 * rx);
    const double vfov = 2.0 * std::atan(pixelSize * 0.5 * ry) * 180.0 / pi;
    const double fx = 1.0 / std::tan((0.5 * c.hfov_deg) * pi / 180.0);
    const double fy = 1.0 / std::tan((0.5 * vfov) * pi / 180.0);
    const double n = 0.1, fr = 10000.0, d = 1.0 / (n - fr);
    const double persp[16] = {fx, 0, 0, 0, 0, -fy, 0, 0, 0, 0, (n + fr) * d, 2.0 * n * fr * d, 0, 0, -1, 0};
    double w2n[16]; Mat4Mul(persp, w2c, w2n);
    double n2w[16]; Mat4Inv(w2n, n2w);
    const double tr[16] = {1, 0, 0, -1, 0, 1, 0, -1, 0, 0, 1, 0, 0, 0, 0, 1};
    const double sc[16] = {2.0 / rx, 0, 0, 0, 0, 2.0 / ry, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1};
    double t1[16]; Mat4Mul(n2w, tr, t1);
    Mat4Mul(t1, sc, S.r2w);
    S.camPos = V3(c.position[0], c.position[1], c.position[2]);
    S.resX = c.res_x; S.resY = c.res_y;
}

static Scene *BuildScene(const pg3g_scene_desc *d) {
    Scene *S = new Scene();
    for (uint32_t i = 0; i < d->num_triangles; ++i) {
        Tri t;
        for (int k = 0; k < 3; ++k) t.p[k] = V3(d->triangles[i * 9 + k * 3], d->triangles[i * 9 + k * 3 + 1], d->triangles[i * 9 + k * 3 + 2]);
        t.n = Normalize(Cross(t.p[1] - t.p[0], t.p[2] - t.p[0]));          // scene_graph.hxx:100
        t.mat = d->triangle_mat[i];
        S->tris.push_back(t);
    }
    for (uint32_t i = 0; i < d->num_spheres; ++i)
        S->sphs.push_back(Sphere{V3(d->spheres[i * 4], d->spheres[i * 4 + 1], d->spheres[i * 4 + 2]), d->spheres[i * 4 + 3], d->sphere_mat[i]});
    for (uint32_t i = 0; i < d->num_materials; ++i) S->mats.push_back(Material::FromDesc(d->materials[i]));
    for (uint32_t i = 0; i < d->num_lights; ++i) {
        const pg3g_light &l = d->lights[i];
        Light L;
        L.type = l.type;
        const V3 pw(l.power[0], l.power[1], l.power[2]);
        if (l.type == PG3G_LIGHT_AREA) L.InitArea(V3(l.p0[0], l.p0[1], l.p0[2]), V3(l.p1[0], l.p1[1], l.p1[2]), V3(l.p2[0], l.p2[1], l.p2[2]), pw);
        else if (l.type == PG3G_LIGHT_POINT) { L.p0 = V3(l.p0[0], l.p0[1], l.p0[2]); L.radiance = pw / (4 * kPiF); }
        else { L.radiance = pw; S->bgLight = (int)i; }
        S->lights.push_back(L);
    }
    if (d->em.width > 0 && d->em.texels) {
        S->em.w = d->em.width; S->em.h = d->em.height; S->em.bilinear = d->em.bilinear != 0;
        S->em.data.resize((size_t)S->em.w * S->em.h);
        for (size_t i = 0; i < S->em.data.size(); ++i) S->em.data[i] = V3(d->em.texels[i * 3], d->em.texels[i * 3 + 1], d->em.texels[i * 3 + 2]);
        S->emSampler = d->em.sampler;
        S->simple.Build(S->em);   // also serves EstimateIrradiance's MC fallback (em.hxx:33-49)
        if (S->emSampler == PG3G_EM_STEERABLE) {
            const uint32_t mn = d->em.st_min_subdiv_level >= 0 ? (uint32_t)d->em.st_min_subdiv_level : 4u;
            const uint32_t mx = d->em.st_max_subdiv_level >= 0 ? (uint32_t)d->em.st_max_subdiv_level : 9u;
            S->steer.Build(S->em, d->em.st_max_approx_error, mn, mx);
        }
    }
    CameraSetup(*S, d->camera);
    S->params = d->params;
    return S;
}

}  // namespace orc

using namespace orc;

extern "C" {

void *oracle_scene_create(const pg3g_scene_desc *d) { return BuildScene(d); }
void oracle_scene_destroy(void *s) { delete static_cast<Scene *>(s); }
void oracle_set_params(void *s, const pg3g_render_params *p) { static_cast<Scene *>(s)->params = *p; }

// Renders iterations [first, first+n) into sum_rgb (FP64 per-pixel sums, res_x*res_y*3, added to).
// Each pixel is independent (counter-based RNG), so pixels are split over OpenMP threads.
// Returns the wall time in seconds.
double oracle_render(void *sp, uint32_t first, uint32_t n, int threads, double *sum_rgb,
                     uint32_t pixel_begin, uint32_t pixel_end) {
    const Scene &S = *static_cast<Scene *>(sp);
    Renderer R(S);
    const uint32_t npix = S.resX * S.resY;
    if (pixel_end == 0 || pixel_end > npix) pixel_end = npix;
    if (threads > 0) omp_set_num_threads(threads);
    const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t pix = pixel_begin; pix < (int64_t)pixel_end; ++pix) {
        const uint32_t x = (uint32_t)pix % S.resX, y = (uint32_t)pix / S.resX;
        double acc[3] = {0, 0, 0};
        for (uint32_t it = first; it < first + n; ++it) {
            PathRng rng{S.params.seed, (uint32_t)pix, it, S.params.scene_id, 0};
            const float jx = rng.Get(0, 0, 0), jy = rng.Get(0, 0, 1);   // path_tracer_base.hxx:74-77
            V3 org, dir;
            S.GenerateRay(float(x) + jx, float(y) + jy, org, dir);
            const V3 L = R.Radiance(org, dir, rng);
            acc[0] += L.x; acc[1] += L.y; acc[2] += L.z;                   // frame_buffer.hxx:30-46
        }
        sum_rgb[pix * 3] += acc[0]; sum_rgb[pix * 3 + 1] += acc[1]; sum_rgb[pix * 3 + 2] += acc[2];
    }
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

// ---- function-level entry points --------------------------------------------------------------
void oracle_bsdf_eval(const pg3g_material *md, uint32_t n, const float *wil, const float *wol, float *f,
                      float *pdf, float *comp) {
    const Material m = Material::FromDesc(*md);
    for (uint32_t i = 0; i < n; ++i) {
        MatRec r;
        r.wil = V3(wil[i * 3], wil[i * 3 + 1], wil[i * 3 + 2]);
        r.wol = V3(wol[i * 3], wol[i * 3 + 1], wol[i * 3 + 2]);
        m.Eval(r);
        f[i * 3] = r.att.x; f[i * 3 + 1] = r.att.y; f[i * 3 + 2] = r.att.z;
        pdf[i] = r.pdfW; comp[i] = r.compProb;
    }
}
void oracle_bsdf_sample(const pg3g_material *md, uint32_t n, const float *wol, const float *u, float *wil,
                        float *f, float *pdf, float *comp) {
    const Material m = Material::FromDesc(*md);
    for (uint32_t i = 0; i < n; ++i) {
        MatRec r;
        r.wol = V3(wol[i * 3], wol[i * 3 + 1], wol[i * 3 + 2]);
        m.Sample(u[i * 3], u[i * 3 + 1], u[i * 3 + 2], r);
        wil[i * 3] = r.wil.x; wil[i * 3 + 1] = r.wil.y; wil[i * 3 + 2] = r.wil.z;
        f[i * 3] = r.att.x; f[i * 3 + 1] = r.att.y; f[i * 3 + 2] = r.att.z;
        pdf[i] = r.pdfW; comp[i] = r.compProb;
    }
}
void oracle_em_sample(void *sp, uint32_t n, const float *normals, const int32_t *flags, const float *u,
                      float *dir, float *pdf, float *radcos, int32_t *ok) {
    const Scene &S = *static_cast<Scene *>(sp);
    Material m;
    for (uint32_t i = 0; i < n; ++i) {
        Frame f; f.SetFromZ(V3(normals[i * 3], normals[i * 3 + 1], normals[i * 3 + 2]));
        m.front = flags[i] & 1; m.back = (flags[i] & 2) != 0;
        LightSample o;
        bool good;
        const float u0 = u[i * 3], u1 = u[i * 3 + 1], u2 = u[i * 3 + 2];
        switch (S.emSampler) {
        case PG3G_EM_SIMPLE: good = S.simple.Sample(f, m.front, m.back, u0, u1, o.wig, o.pdfW, o.sample); break;
        case PG3G_EM_STEERABLE: good = S.steer.Sample(f, m.front, m.back, u2, u0, u1, o.wig, o.pdfW, o.sample); break;
        default: {
            const V3 w = SampleCosSphereParamPdfW(u0, u1, u2, m.front, m.back, o.pdfW);
            o.wig = f.ToWorld(w);
            o.sample = S.em.Evaluate(o.wig) * std::abs(w.z);
            good = true;
        }
        }
        ok[i] = good ? 1 : 0;
        dir[i * 3] = o.wig.x; dir[i * 3 + 1] = o.wig.y; dir[i * 3 + 2] = o.wig.z;
        pdf[i] = o.pdfW;
        radcos[i * 3] = o.sample.x; radcos[i * 3 + 1] = o.sample.y; radcos[i * 3 + 2] = o.sample.z;
    }
}
void oracle_em_pdf(void *sp, uint32_t n, const float *dirs, const float *normals, const int32_t *flags, float *pdf) {
    const Scene &S = *static_cast<Scene *>(sp);
    for (uint32_t i = 0; i < n; ++i) {
        Frame f; f.SetFromZ(V3(normals[i * 3], normals[i * 3 + 1], normals[i * 3 + 2]));
        pdf[i] = S.EmPdf(V3(dirs[i * 3], dirs[i * 3 + 1], dirs[i * 3 + 2]), f, flags[i] & 1, (flags[i] & 2) != 0);
    }
}
void oracle_em_eval(void *sp, uint32_t n, const float *dirs, float *rgb) {
    const Scene &S = *static_cast<Scene *>(sp);
    for (uint32_t i = 0; i < n; ++i) {
        const V3 v = S.em.Evaluate(V3(dirs[i * 3], dirs[i * 3 + 1], dirs[i * 3 + 2]));
        rgb[i * 3] = v.x; rgb[i * 3 + 1] = v.y; rgb[i * 3 + 2] = v.z;
    }
}
void oracle_em_eval_uv(void *sp, uint32_t n, const float *uv, float *rgb) {
    const Scene &S = *static_cast<Scene *>(sp);
    for (uint32_t i = 0; i < n; ++i) {
        const V3 v = S.em.Evaluate(uv[i * 2], uv[i * 2 + 1]);
        rgb[i * 3] = v.x; rgb[i * 3 + 1] = v.y; rgb[i * 3 + 2] = v.z;
    }
}
// The light-picking estimate of light `li` (AreaLight::EstimateContribution, lights.hxx:133-167;
// BackgroundLight / EnvironmentMap::EstimateIrradiance, lights.hxx:451-467, em.hxx:123-170) at n
// surface points with normals and side flags (bit 0 front, bit 1 back).  Point i draws its random
// numbers from pixel i, iteration `iteration`, vertex 1 of the scene's seed.
void oracle_estimate_contribution(void *sp, int32_t li, uint32_t n, const float *pts, const float *normals,
                                  const int32_t *flags, uint32_t iteration, float *out) {
    const Scene &S = *static_cast<Scene *>(sp);
    Material m;
    for (uint32_t i = 0; i < n; ++i) {
        Frame f; f.SetFromZ(V3(normals[i * 3], normals[i * 3 + 1], normals[i * 3 + 2]));
        m.front = flags[i] & 1; m.back = (flags[i] & 2) != 0;
        PathRng rng{S.params.seed, i, iteration, 0, 0};
        out[i] = S.EstimateContribution(li, V3(pts[i * 3], pts[i * 3 + 1], pts[i * 3 + 2]), f, m, rng, 1);
    }
}
float oracle_em_avg_pixel_lum(void *sp, uint32_t x, uint32_t y) { return static_cast<Scene *>(sp)->em.AvgPixelLum(x, y); }
void oracle_camera_ray(void *sp, float x, float y, float *org, float *dir) {
    V3 o, d; static_cast<Scene *>(sp)->GenerateRay(x, y, o, d);
    org[0] = o.x; org[1] = o.y; org[2] = o.z; dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
}
// Scene::Intersect / Occluded of n rays (scene.hxx:42-74): dist -1 and material -1 for a miss.
void oracle_intersect_batch(void *sp, uint32_t n, const float *org, const float *dir, const float *tmin, float *dist,
                            float *normal, int32_t *mat) {
    const Scene &S = *static_cast<Scene *>(sp);
    for (uint32_t i = 0; i < n; ++i) {
        Isect r; r.dist = 1e36f;
        const bool hit = S.Intersect(V3(org[3 * i], org[3 * i + 1], org[3 * i + 2]),
                                     V3(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]), tmin[i], r);
        dist[i] = hit ? r.dist : -1.f;
        mat[i] = hit ? r.matID : -1;
        normal[3 * i] = hit ? r.normal.x : 0.f;
        normal[3 * i + 1] = hit ? r.normal.y : 0.f;
        normal[3 * i + 2] = hit ? r.normal.z : 0.f;
    }
}
void oracle_occluded_batch(void *sp, uint32_t n, const float *p, const float *dir, const float *tmax, int32_t *occ) {
    const Scene &S = *static_cast<Scene *>(sp);
    for (uint32_t i = 0; i < n; ++i)
        occ[i] = S.Occluded(V3(p[3 * i], p[3 * i + 1], p[3 * i + 2]), V3(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]),
                            tmax[i]) ? 1 : 0;
}

// ---- primitives for known-answer tests ----------------------------------------------------------
void oracle_sh_basis(const float *dir, float *out9) { const SH9 b = ShBasis(V3(dir[0], dir[1], dir[2])); for (int i = 0; i < 9; ++i) out9[i] = b[i]; }
void oracle_clamped_cos(const float *n, int comp, float *out9) { const SH9 c = ClampedCosCoeffs(V3(n[0], n[1], n[2]), comp != 0); for (int i = 0; i < 9; ++i) out9[i] = c[i]; }
float oracle_fresnel_dielectric(float c, float eta) { return FresnelDielectric(c, eta); }
float oracle_fresnel_conductor(float c, float eta, float k) { return FresnelConductor(c, eta, k); }
float oracle_ggx_d(const float *h, float a) { return DistributionGgx(V3(h[0], h[1], h[2]), a); }
float oracle_ggx_g1(const float *v, const float *h, float a) { return SmithG1(V3(v[0], v[1], v[2]), V3(h[0], h[1], h[2]), a); }
void oracle_ggx_vndf_sample(const float *wo, float a, float u0, float u1, float *h) {
    const V3 r = SampleGgxVisibleNormals(V3(wo[0], wo[1], wo[2]), a, u0, u1); h[0] = r.x; h[1] = r.y; h[2] = r.z;
}
float oracle_ggx_vndf_pdf(const float *wo, const float *h, float a) {
    const V3 H(h[0], h[1], h[2]);
    return GgxVndfPdf(V3(wo[0], wo[1], wo[2]), H, DistributionGgx(H, a), a);
}
void oracle_halfway_refraction(const float *wi, const float *wo, float eta, float *h) {
    const V3 r = HalfwayRefraction(V3(wi[0], wi[1], wi[2]), V3(wo[0], wo[1], wo[2]), eta); h[0] = r.x; h[1] = r.y; h[2] = r.z;
}
void oracle_refract(const float *in, const float *n, float eta, float *out) {
    V3 o; bool a; Refract(o, a, V3(in[0], in[1], in[2]), V3(n[0], n[1], n[2]), eta); out[0] = o.x; out[1] = o.y; out[2] = o.z;
}
void oracle_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint64_t seed, uint32_t *out4) { Philox::Gen(out4, c0, c1, c2, c3, seed); }
void oracle_latlong2dir(float u, float v, float *d) { const V3 r = LatLong2Dir(u, v); d[0] = r.x; d[1] = r.y; d[2] = r.z; }
void oracle_dir2latlong(const float *d, float *uv) { const V2 r = Dir2LatLong(V3(d[0], d[1], d[2])); uv[0] = r.x; uv[1] = r.y; }
void oracle_sample_tri_bilinear(float ux, float uy, float a, float b, float c, float *bary2) {
    SteerableSampler::SampleTriBilinear(ux, uy, a, b, c, bary2[0], bary2[1]);
}
// s and t of SampleTriangleBilinear (em_steerable_sampler.hxx:3709-3747, 3882-3905)
void oracle_sample_tri_st(float ux, float uy, float a, float b, float c, float *st) {
    SteerableSampler::SampleTriST(ux, uy, a, b, c, st[0], st[1]);
}
// 1-D distribution: builds from f (n), samples m uniforms
void oracle_dist1d(const float *f, uint32_t n, uint32_t m, const float *u, float *x, uint32_t *seg, float *pdf,
                   float *cdf_out, float *cdf_f32_out) {
    Dist1D d; d.Build(f, n);
    for (uint32_t i = 0; i <= n; ++i) cdf_out[i] = d.cdf[i];
    if (cdf_f32_out) { std::vector<float> c; float I; Dist1D::ComputeCdfF32(c, I, f, n); for (uint32_t i = 0; i <= n; ++i) cdf_f32_out[i] = c[i]; }
    for (uint32_t i = 0; i < m; ++i) { size_t s; d.Sample(u[i], x[i], s, pdf[i]); seg[i] = (uint32_t)s; }
}
// steerable introspection
void oracle_steerable_info(void *sp, uint32_t *leaves, uint32_t *verts) {
    const Scene &S = *static_cast<Scene *>(sp);
    std::vector<const SteerableSampler::Node *> L; S.steer.Leaves(L);
    *leaves = (uint32_t)L.size(); *verts = (uint32_t)S.steer.verts.size();
}
void oracle_steerable_leaves(void *sp, float *dirs) {
    const Scene &S = *static_cast<Scene *>(sp);
    std::vector<const SteerableSampler::Node *> L; S.steer.Leaves(L);
    for (size_t i = 0; i < L.size(); ++i)
        for (int k = 0; k < 3; ++k) { const V3 &d = S.steer.verts[L[i]->tri.vi[k]].dir; dirs[i * 9 + k * 3] = d.x; dirs[i * 9 + k * 3 + 1] = d.y; dirs[i * 9 + k * 3 + 2] = d.z; }
}
void oracle_steerable_root_weight(void *sp, float *w9) {
    const Scene &S = *static_cast<Scene *>(sp);
    const SH9 w = S.steer.root->leaf ? S.steer.TriWeight(S.steer.root->tri) : SteerableSampler::NodeWeight(*S.steer.root);
    for (int i = 0; i < 9; ++i) w9[i] = w[i];
}
// Parts of the steerable sampler's pdf for the clamped-cosine lobe of `normal`: GetBasePdf of dir
// and of -dir and ComputeWholeIntegral (em_steerable_sampler.hxx:2387-2432), per direction.
void oracle_steerable_pdf_parts(void *sp, const float *normal, uint32_t m, const float *dirs, float *out3) {
    const Scene &S = *static_cast<Scene *>(sp);
    const SH9 c = ClampedCosCoeffs(V3(normal[0], normal[1], normal[2]), true);
    const float whole = S.steer.WholeIntegral(c);
    for (uint32_t i = 0; i < m; ++i) {
        const V3 d(dirs[i * 3], dirs[i * 3 + 1], dirs[i * 3 + 2]);
        out3[i * 3] = S.steer.BasePdf(d, c);
        out3[i * 3 + 1] = S.steer.BasePdf(-d, c);
        out3[i * 3 + 2] = whole;
    }
}
// Samples triangle leaves via PickTriangle for a normal; returns leaf index per uniform (for the
// picking-frequency test, em_steerable_sampler.hxx:4778-4925)
void oracle_steerable_pick(void *sp, const float *normal, uint32_t m, const float *u, uint32_t *leaf_index,
                           float *leaf_integral_rel) {
    const Scene &S = *static_cast<Scene *>(sp);
    std::vector<const SteerableSampler::Node *> L; S.steer.Leaves(L);
    const SH9 c = ClampedCosCoeffs(V3(normal[0], normal[1], normal[2]), true);
    const float whole = S.steer.WholeIntegral(c);
    for (size_t i = 0; i < L.size() && leaf_integral_rel; ++i) leaf_integral_rel[i] = Dot9(S.steer.TriWeight(L[i]->tri), c) / whole;
    for (uint32_t i = 0; i < m; ++i) {
        float uu = u[i];
        const SteerableSampler::Node *n = S.steer.PickTriangle(c, uu);
        leaf_index[i] = (uint32_t)(std::find(L.begin(), L.end(), n) - L.begin());
    }
}

}  // extern "C"
