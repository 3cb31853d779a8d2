// Per-sample radiance estimators of the reference renderers, written as iterative device code.
// PathTracer::EstimateIncomingRadiancePT (path_tracer.hxx:167-430) with splitting budget 1 and no
// indirect clipping is a chain: pixel = emitted(1) + sum_v B_v * (NEE_v + E_v), where B_v is the
// path throughput up to vertex v, NEE_v the MIS-weighted light sample (path_tracer_base.hxx:
// 350-425) and E_v the MIS-weighted emission found by the BSDF sample of v (path_tracer.hxx:
// 327-369); the sums accumulate forward (ChainPath, RadiancePT).  The indirect clipping of
// path_tracer.hxx:400-402 needs each vertex's whole suffix: with splitting at the first vertex
// only (every integer -sb at splitting level 1, the reference's defaults) SplitPath folds per-lane
// records of each chain back to front; every other splitting configuration takes the general
// recursion walker RadiancePTTree.
#pragma once

#include <type_traits>
#include "dev_scene.hpp"

namespace pg3 {

constexpr uint32_t kMaxPickLights = 8;    // per-vertex light-contribution cache kept in registers

constexpr int kEmDynamic = -2;   // kernel template value: light spec read from the scene at run time

// Sum over the 64 lanes of a wavefront.
__device__ __forceinline__ uint32_t WaveSum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Work of one lane, summed per wavefront into a launch's counters (FlushTally): shaded path
// vertices (segments), camera samples, and the rays actually traced -- extension rays
// (Scene::Intersect, scene.hxx:42-59) and shadow rays (Scene::Occluded, :61-74).  The w* fields
// are wavefront totals, only ever changed in wave-uniform code (they stay in scalar registers;
// the persistent item loop counts samples and extension rays from its ballots).
struct RayTally {
    uint32_t segments = 0, samples = 0, ext = 0, shadow = 0;
    uint32_t wSamples = 0, wExt = 0, wShadow = 0;
};
// counters: [0] segments, [1] camera samples, [2] extension rays, [3] shadow rays (64-bit).
__device__ __forceinline__ void FlushTally(const RayTally &t, unsigned long long *counters) {
    const uint32_t a = WaveSum(t.segments), b = WaveSum(t.samples) + t.wSamples, c = WaveSum(t.ext) + t.wExt,
                   d = WaveSum(t.shadow) + t.wShadow;
    if (__lane_id() == 0 && counters) {
        if (a) atomicAdd(counters + 0, (unsigned long long)a);
        if (b) atomicAdd(counters + 1, (unsigned long long)b);
        if (c) atomicAdd(counters + 2, (unsigned long long)c);
        if (d) atomicAdd(counters + 3, (unsigned long long)d);
    }
}

// MISWeight2Power, path_tracer_base.hxx:522-541 (sample counts 1)
PG3_DEV float MisPower(float p1, float p2) {
    const float q1 = p1 * p1, q2 = p2 * p2;
    return FastDiv(q1, q1 + q2);
}

struct LightCache {            // LightSamplingContext, path_tracer_base.hxx:22-32
    float est[kMaxPickLights];
    float sum;
};
PG3_DEV void FillLightCache(const SceneView &S, const LightSpec &spec, F3 sp, const Frame &fr, uint32_t sides,
                            const RngKey &key, uint32_t vtx, LightCache &lc) {
    lc.sum = 0.f;
#pragma unroll
    for (uint32_t k = 0; k < kMaxPickLights; ++k) lc.est[k] = 0.f;
    // one copy of EstimateContribution in the code (a rolled loop); the table stays in registers
#pragma unroll 1
    for (uint32_t i = 0; i < spec.numLights; ++i) {
        const float e = EstimateContribution(S, spec, i, sp, fr, sides, key, vtx);
#pragma unroll
        for (uint32_t k = 0; k < kMaxPickLights; ++k)
            if (k == i) lc.est[k] = e;
        lc.sum += e;
    }
}
// LightPickingProbability, path_tracer_base.hxx:260-313
PG3_DEV float PickProb(const LightSpec &spec, const LightCache &lc, uint32_t id) {
    if (spec.numLights == 0) return 0.f;
    if (spec.numLights == 1) return 1.f;
    float e = 0.f;
#pragma unroll
    for (uint32_t i = 0; i < kMaxPickLights; ++i) if (i == id) e = lc.est[i];
    return (lc.sum > 0.f) ? e / lc.sum : 1.f / spec.numLights;
}
// PickSingleLight, path_tracer_base.hxx:191-257
PG3_DEV int32_t PickLight(const LightSpec &spec, const LightCache &lc, float u, float &prob) {
    const uint32_t n = spec.numLights;
    if (n == 0) return -1;
    if (n == 1) { prob = 1.f; return 0; }
    if (lc.sum > 0.f) {
        const float rv = u * lc.sum;
        float cdf = 0.f, prev = 0.f;
        uint32_t li = 0;
        bool found = false;
#pragma unroll
        for (uint32_t i = 0; i < kMaxPickLights; ++i) {
            if (i < n && !found) {
                prev = cdf;
                cdf += lc.est[i];
                li = i;
                if (!(rv > cdf)) found = true;
            }
        }
        // rv <= sum, so the scan always stops inside the table; keep the index in range regardless
        if (!found) li = n - 1;
        prob = (cdf - prev) / lc.sum;
        return (int32_t)li;
    }
    prob = 1.f / n;
    return min((int32_t)(u * n), (int32_t)(n - 1));
}

// Tells the compiler which material types a class kernel meets: the type becomes one of the
// class's own constants, so the BSDF code of every other type is dead code in that kernel.
template <int kCls>
PG3_DEV void NarrowMaterial(DevMaterial &m) {
    if constexpr (kCls == kClsDiffuse) m.type = (m.type == kMatPhong) ? kMatPhong : kMatLambert;
    else if constexpr (kCls == kClsGgxCond) m.type = kMatGgxCond;
    else if constexpr (kCls == kClsGgxDiel) m.type = kMatGgxDiel;
    else if constexpr (kCls == kClsDirac) m.type = (m.type == kMatSmoothCond) ? kMatSmoothCond : kMatSmoothDiel;
}

// State of one path between two vertices of PathTracer::EstimateIncomingRadiancePT (-sb 1).
struct PathState {
    F3 org, dir;
    float tmin;
    F3 beta;             // throughput up to the current vertex (product of indirect factors)
    F3 eThr;             // weight of emission found by the last BSDF sample (MIS applied on hit)
    F3 L;                // accumulated radiance estimate of the sample
    float prevBsdfPdf;   // pdfW * compProb of the last BSDF sample
    F3 prevN;            // shading normal of the previous vertex (EM pdf w.r.t. its frame)
    uint32_t len;        // path length of the next vertex (1 = camera ray hit)
    uint32_t flags;      // kHasPrev | kPrevFinite | kComputeRefl | previous sides << 4
    float lcSum;         // sum of the previous vertex's light-contribution cache (> 1 light)
};
enum : uint32_t { kHasPrev = 1u, kPrevFinite = 2u, kComputeRefl = 4u };

PG3_DEV void InitPath(PathState &ps, F3 org, F3 dir) {
    ps.org = org;
    ps.dir = dir;
    ps.tmin = 0.f;
    ps.beta = mk(1.f);
    ps.eThr = mk(0.f);
    ps.L = mk(0.f);
    ps.prevBsdfPdf = 0.f;
    ps.prevN = mk(0.f, 0.f, 1.f);
    ps.len = 1;
    ps.flags = kComputeRefl | kPrevFinite;
    ps.lcSum = 0.f;
}

// LightPickingProbability of light `id` at the previous vertex of the path (path_tracer_base.hxx:
// 260-313).  Only the cache's sum travels with the path; the one entry needed is recomputed from
// the previous vertex (position ps.org, normal ps.prevN, sides, RNG vertex ps.len - 1), which
// gives the cached value bit for bit (same function, same arguments).
PG3_DEV float PrevPickProb(const SceneView &S, const LightSpec &spec, const RngKey &key, const PathState &ps,
                           uint32_t id) {
    if (spec.numLights == 0) return 0.f;
    if (spec.numLights == 1) return 1.f;
    if (!(ps.lcSum > 0.f)) return 1.f / spec.numLights;
    const float e = EstimateContribution(S, spec, id, ps.org, FrameFromZ(ps.prevN), ps.flags >> 4, key, ps.len - 1u);
    return e / ps.lcSum;
}

// Next-event request produced at a vertex: the contribution is added when the segment from `p`
// along `dir` up to `tmax` is unoccluded (Scene::Occluded, scene.hxx:61-74).
struct ShadowReq {
    F3 p, dir, contrib;
    float tmax;
    bool valid;
};

// Emission of an area light hit by the previous BSDF sample (path_tracer.hxx:198-210, MIS as
// :327-369) and the checks that end the path at this vertex (:212-238).  Returns true if the
// vertex is shaded.
PG3_DEV bool HitEmissionAndChecks(const SceneView &S, const LightSpec &spec, const RngKey &key, PathState &ps, F3 sp,
                                  F3 wol, const DevMaterial &mat, uint32_t &segments) {
    const uint32_t len = ps.len;
    const bool hasPrev = ps.flags & kHasPrev, prevFinite = ps.flags & kPrevFinite;
    if (mat.lightId >= 0 && len >= S.minPL) {
        const DevLight Lt = LoadGlobal(S.lights + mat.lightId);
        F3 Le = mk(0.f);
        if (Lt.type == kLightArea && wol.z > 0.f) Le = Lt.radiance;
        if (!hasPrev) {
            ps.L += Le;
        } else if (!IsZero(Le)) {
            if (prevFinite) {
                const float lp = AreaEmissionPdf(Lt, sp, ps.org) * PrevPickProb(S, spec, key, ps, (uint32_t)mat.lightId);
                ps.L += ps.eThr * Le * MisPower(ps.prevBsdfPdf, lp);
            } else {
                ps.L += ps.eThr * Le;
            }
        }
    }
    if (!(ps.flags & kComputeRefl)) return false;
    if (ReflectanceZero(mat)) return false;
    if (S.maxPL > 0 && len >= S.maxPL) return false;
    if (len >= 1200u) return false;
    ++segments;
    return true;
}

// Shadow-ray request of a light sample (AddMISLightSampleContribution, path_tracer_base.hxx:
// 350-425): the contribution is added if the segment to the light is unoccluded.
PG3_DEV void LightSampleRequest(const PathState &ps, const DevMaterial &mat, const Frame &fr, F3 wol, F3 sp,
                                const LightSample &ls, ShadowReq &sr) {
    if (MaxC(ls.sample) > 0.f) {
        BsdfRec r;
        EvalBsdf(mat, ToLocal(fr, ls.wig), wol, r);
        if (ls.pdfW != kInf) {
            const float lightPdf = ls.pdfW * ls.prob;
            const float bsdfPdf = r.pdfW * r.compProb;
            sr.contrib = ps.beta * FastDiv(ls.sample * r.att * MisPower(lightPdf, bsdfPdf), lightPdf);
        } else {
            sr.contrib = ps.beta * FastDiv(ls.sample * r.att, ls.prob);
        }
        sr.p = sp;
        sr.dir = ls.wig;
        sr.tmax = ls.dist;
        sr.valid = true;
    }
}

// BSDF sample, Russian roulette and the path update (path_tracer.hxx:267-303).  Returns true if
// the path continues.
PG3_DEV bool ContinuePath(const SceneView &S, PathState &ps, F3 sp, const Frame &fr, F3 wol, const DevMaterial &mat,
                          const VertexRng &vr, float lcSum) {
    float rr = 1.0f;
    bool cut = false;
    if (S.maxPL == 0) {
        rr = Clampf(RRProb(mat, wol), 0.0f, 0.997f);
        cut = vr.u[kDimRR] > rr;
    }
    BsdfRec r;
    SampleBsdf(mat, wol, vr.u[kDimBsdf0], vr.u[kDimBsdf1], vr.u[kDimBsdf2], r);
    if (IsBlocker(r)) return false;
    const float cosAbs = fabsf(r.wil.z);
    const bool finite = (r.pdfW != kInf);
    if (finite) {
        ps.prevBsdfPdf = r.pdfW * r.compProb;
        ps.eThr = FastDiv(ps.beta * r.att * cosAbs, ps.prevBsdfPdf);
        if (!cut) ps.beta = ps.beta * FastDiv(r.att * cosAbs, r.pdfW * 1u * rr * r.compProb);
    } else {
        ps.eThr = ps.beta * r.att / (1.f * r.compProb);
        if (!cut) ps.beta = ps.beta * (r.att / (1u * rr * r.compProb));
    }
    ps.flags = kHasPrev | (finite ? kPrevFinite : 0u) | (cut ? 0u : kComputeRefl) | (mat.sides << 4);
    ps.prevN = fr.z;
    ps.lcSum = lcSum;
    ps.org = sp;
    ps.dir = ToWorld(fr, r.wil);
    ps.tmin = EpsRayCos(cosAbs);
    ps.len = ps.len + 1;
    return true;
}

// The EM lookups of a vertex in a scene with a steerable EM.  An escaped path (emission
// with MIS, path_tracer.hxx:408-422: the EM pdf of its direction at the previous vertex) and a light
// sample of the EM (SampleImpl and SampleHemisphereImpl, em_steerable_sampler.hxx:2084-2217) need
// the same lookups: the hemisphere's clamped-cosine coefficients and integrals, the EM radiance at
// one direction and the luminance at its antipode (one lat-long mapping), and the two GetBasePdf
// terms (:2387-2413) of that direction and its antipode.  Lanes of both kinds run these lookups in
// one pass, so a wavefront does not issue them once for its escaped lanes and again for its
// shading lanes.  A light-sample lane (nee) samples around the normal qn of its vertex (sides
// qsides, random numbers uSide, u0, u1, pick probability prob) and gets ls / lsOk; an escaped lane
// gets the radiance Le along qdir and its pdf emPdf under the sampling of the previous vertex
// (normal qn, sides qsides).
struct EmQuery {
    F3 Le;
    float emPdf;
    LightSample ls;
    bool lsOk;
};
PG3_DEV void SteerableEmQuery(const DevEm &E, const LightSpec &spec, bool nee, F3 qn, F3 qdir, uint32_t qsides,
                              float uSide, float u0, float u1, float prob, EmQuery &q) {
    const bool front = qsides & kSideFront, back = qsides & kSideBack;
    bool ok = front || back;
    float ui = 0.f, li = 0.f, wi = 0.f;
    bool up = front;
    if (front && back) {
        ui = WholeIntegral(E, ClampedCos(qn));
        li = WholeIntegral(E, ClampedCos(-qn));
        wi = ui + li;
        if (IsTiny(wi)) ok = false;
        up = nee ? (uSide * wi < ui) : (Dot(qdir, qn) >= 0.f);
    }
    const F3 hn = up ? qn : -qn;
    const SH9 c = ClampedCos(hn);
    const float whole = WholeIntegral(E, c);
    float triPdf = 0.f, cs = 0.f;
    bool flip = false;
    if (nee && ok) {
        float val;
        const F3 sdir = SteerTriangleSample(E, c, u0, u1, spec.tree, val);
        triPdf = IsTiny(whole) ? 0.f : FastDiv(val, whole);
        cs = Dot(sdir, hn);
        flip = cs < 0.f;
        qdir = sdir;
        if (flip) { qdir = -sdir; cs *= -1.f; }
    }
    float u, v, uo, vo;
    Dir2LatLong(qdir, u, v);
    q.Le = EmEvalUV(E, u, v);
    Antipode(u, v, uo, vo);
    const float lumOpp = EmLumUV(E, uo, vo);
    // GetBasePdf of qdir and of -qdir; a light sample needs the one of -(sampled direction)
    const float t1 = BasePdfLum(E, qdir, Luminance(q.Le), c);
    const float t2 = BasePdfLum(E, -qdir, lumOpp, c);
    q.emPdf = 0.f;
    q.lsOk = false;
    if (nee) {
        q.lsOk = ok;
        float pdf = triPdf;
        pdf += flip ? t1 : t2;
        if (front && back) pdf *= FastDiv(up ? ui : li, wi);
        q.ls.wig = qdir;
        q.ls.pdfW = pdf;
        q.ls.sample = q.Le * cs;
        q.ls.prob = prob;
        q.ls.dist = 3.402823466e+38f;   // em_sampler.hxx:57
    } else if (ok) {
        const float p = (Dot(qdir, hn) < 0.f) ? 0.f : t1 + t2;
        q.emPdf = (front && back) ? p * FastDiv(up ? ui : li, wi) : p;
    }
}

// Vertex of the path tracer with a steerable EM; the same results as the generic vertex below, in
// another order (the EM lookups of SteerableEmQuery in one pass).
template <int kCls = kClsAny>
PG3_DEV bool ShadeVertexSteerable(const SceneView &S, const LightSpec &spec, const RngKey &key, PathState &ps,
                                  bool hit, const Hit &h, ShadowReq &sr, uint32_t &segments) {
    const uint32_t len = ps.len;
    F3 sp = mk(0.f), wol = mk(0.f);
    Frame fr{};
    DevMaterial mat{};
    VertexRng vr{};
    LightCache lc;
    lc.sum = 0.f;
    bool shade = false;            // hit lane that shades this vertex
    bool emq = false, nee = false;  // lane runs the EM lookups; they serve a light sample
    F3 qn = ps.prevN;
    uint32_t qsides = ps.flags >> 4;
    float uSide = 0.f, u0 = 0.f, u1 = 0.f, prob = 1.f;
    if (hit) {
        sp = ps.org + ps.dir * h.dist;
        fr = FrameFromZ(h.n);
        wol = ToLocal(fr, -ps.dir);
        mat = LoadGlobal(S.mats + h.mat);
        NarrowMaterial<kCls>(mat);
        shade = HitEmissionAndChecks(S, spec, key, ps, sp, wol, mat, segments);
        if (shade) {
            RngVertex(key, len, vr);
            if (spec.numLights > 1) FillLightCache(S, spec, sp, fr, mat.sides, key, len, lc);
            if ((len + 1) >= S.minPL) {                                               // :252-265
                const int32_t li = PickLight(spec, lc, vr.u[kDimLightPick], prob);
                // the EM belongs to the background light, so a scene with one light has no other
                if (li >= 0 && (spec.numLights == 1 || li == S.bgLight)) {
                    emq = nee = true;
                    qn = fr.z;
                    qsides = mat.sides;
                    u0 = vr.u[kDimLight0];
                    u1 = vr.u[kDimLight1];
                    uSide = vr.u[kDimLight2];
                } else if (li >= 0) {
                    LightSample ls;
                    if (SampleIllumination(S, spec, (uint32_t)li, sp, fr, mat.sides, vr.u[kDimLight0],
                                           vr.u[kDimLight1], vr.u[kDimLight2], ls)) {
                        ls.prob = prob;
                        LightSampleRequest(ps, mat, fr, wol, sp, ls, sr);
                    }
                }
            }
        }
    } else {
        emq = len >= S.minPL && S.bgLight >= 0;
    }
    EmQuery q;
    q.Le = mk(0.f);
    q.emPdf = 0.f;
    q.lsOk = false;
    if (emq) SteerableEmQuery(S.em, spec, nee, qn, ps.dir, qsides, uSide, u0, u1, prob, q);
    if (!hit) {                                                                       // :408-422
        if (emq) {
            if (!(ps.flags & kHasPrev)) {
                ps.L += q.Le;
            } else if (!IsZero(q.Le)) {
                if (ps.flags & kPrevFinite) {
                    const float lp = q.emPdf * PrevPickProb(S, spec, key, ps, (uint32_t)S.bgLight);
                    ps.L += ps.eThr * q.Le * MisPower(ps.prevBsdfPdf, lp);
                } else {
                    ps.L += ps.eThr * q.Le;
                }
            }
        }
        return false;
    }
    if (!shade) return false;
    // The material is loaded again rather than kept in registers through the EM lookups (the
    // barrier keeps the compiler from reusing the first load).
    __asm__ volatile("" ::: "memory");
    mat = LoadGlobal(S.mats + h.mat);
    NarrowMaterial<kCls>(mat);
    if (q.lsOk) LightSampleRequest(ps, mat, fr, wol, sp, q.ls, sr);
    return ContinuePath(S, ps, sp, fr, wol, mat, vr, lc.sum);
}

// One vertex of EstimateIncomingRadiancePT (path_tracer.hxx:167-430) with one light sample and
// one BSDF sample: adds the MIS-weighted emission found by the previous BSDF sample, emits the
// light-sample request and samples the BSDF.  Returns true if the path continues.
// kMergeEm: kernels whose only light is a steerable EM take ShadeVertexSteerable (with several
// lights its extra live state costs more register spill than it saves).
// kCls: the material class of every hit this call shades (kClsAny: any).
template <bool kMergeEm = false, int kCls = kClsAny>
PG3_DEV bool ShadeVertex(const SceneView &S, const LightSpec &spec, const RngKey &key, PathState &ps, bool hit, const Hit &h,
                         ShadowReq &sr, uint32_t &segments) {
    sr.valid = false;
    if constexpr (kMergeEm) return ShadeVertexSteerable<kCls>(S, spec, key, ps, hit, h, sr, segments);
    const uint32_t len = ps.len;
    const bool hasPrev = ps.flags & kHasPrev, prevFinite = ps.flags & kPrevFinite;
    if (!hit) {
        if (len >= S.minPL && S.bgLight >= 0) {                                       // :408-422
            // the steerable pdf needs the EM at the direction and at its antipode: one mapping
            const bool pair = (spec.em == kEmSteerable);
            EmPair le;
            if (pair) le = EmEvalPair(S.em, ps.dir);
            const F3 Le = pair ? le.at : BgRadiance(S, spec, ps.dir);
            if (!hasPrev) {
                ps.L += Le;
            } else if (!IsZero(Le)) {
                if (prevFinite) {
                    const float lp = BgPdf(S, spec, ps.dir, FrameFromZ(ps.prevN), ps.flags >> 4, pair ? &le : nullptr) *
                                     PrevPickProb(S, spec, key, ps, (uint32_t)S.bgLight);
                    ps.L += ps.eThr * Le * MisPower(ps.prevBsdfPdf, lp);
                } else {
                    ps.L += ps.eThr * Le;
                }
            }
        }
        return false;
    }
    const F3 sp = ps.org + ps.dir * h.dist;
    const Frame fr = FrameFromZ(h.n);
    const F3 wol = ToLocal(fr, -ps.dir);
    DevMaterial mat = LoadGlobal(S.mats + h.mat);
    NarrowMaterial<kCls>(mat);
    if (mat.lightId >= 0 && len >= S.minPL) {                                         // :198-210
        const DevLight Lt = LoadGlobal(S.lights + mat.lightId);
        F3 Le = mk(0.f);
        if (Lt.type == kLightArea && wol.z > 0.f) Le = Lt.radiance;
        if (!hasPrev) {
            ps.L += Le;
        } else if (!IsZero(Le)) {                                                     // :327-369
            if (prevFinite) {
                const float lp = AreaEmissionPdf(Lt, sp, ps.org) * PrevPickProb(S, spec, key, ps, (uint32_t)mat.lightId);
                ps.L += ps.eThr * Le * MisPower(ps.prevBsdfPdf, lp);
            } else {
                ps.L += ps.eThr * Le;
            }
        }
    }
    if (!(ps.flags & kComputeRefl)) return false;                                     // :212-238
    if (ReflectanceZero(mat)) return false;
    if (S.maxPL > 0 && len >= S.maxPL) return false;
    if (len >= 1200u) return false;
    ++segments;

    VertexRng vr;
    RngVertex(key, len, vr);
    const uint32_t sides = mat.sides;
    LightCache lc;
    lc.sum = 0.f;
    if (spec.numLights > 1) FillLightCache(S, spec, sp, fr, sides, key, len, lc);
    if ((len + 1) >= S.minPL) {                                                       // :252-265
        float prob = 1.f;
        const int32_t li = PickLight(spec, lc, vr.u[kDimLightPick], prob);
        LightSample ls;
        if (li >= 0 && SampleIllumination(S, spec, (uint32_t)li, sp, fr, sides, vr.u[kDimLight0], vr.u[kDimLight1],
                                          vr.u[kDimLight2], ls)) {
            ls.prob = prob;
            if (MaxC(ls.sample) > 0.f) {                                              // :350-425
                BsdfRec r;
                EvalBsdf(mat, ToLocal(fr, ls.wig), wol, r);
                if (ls.pdfW != kInf) {
                    const float lightPdf = ls.pdfW * ls.prob;
                    const float bsdfPdf = r.pdfW * r.compProb;
                    sr.contrib = ps.beta * FastDiv(ls.sample * r.att * MisPower(lightPdf, bsdfPdf), lightPdf);
                } else {
                    sr.contrib = ps.beta * FastDiv(ls.sample * r.att, ls.prob);
                }
                sr.p = sp;
                sr.dir = ls.wig;
                sr.tmax = ls.dist;
                sr.valid = true;
            }
        }
    }
    float rr = 1.0f;                                                                  // :267-303
    bool cut = false;
    if (S.maxPL == 0) {
        rr = Clampf(RRProb(mat, wol), 0.0f, 0.997f);
        cut = vr.u[kDimRR] > rr;
    }
    BsdfRec r;
    SampleBsdf(mat, wol, vr.u[kDimBsdf0], vr.u[kDimBsdf1], vr.u[kDimBsdf2], r);
    if (IsBlocker(r)) return false;
    const float cosAbs = fabsf(r.wil.z);
    const bool finite = (r.pdfW != kInf);
    if (finite) {
        ps.prevBsdfPdf = r.pdfW * r.compProb;
        ps.eThr = FastDiv(ps.beta * r.att * cosAbs, ps.prevBsdfPdf);
        if (!cut) ps.beta = ps.beta * FastDiv(r.att * cosAbs, r.pdfW * 1u * rr * r.compProb);
    } else {
        ps.eThr = ps.beta * r.att / (1.f * r.compProb);
        if (!cut) ps.beta = ps.beta * (r.att / (1u * rr * r.compProb));
    }
    ps.flags = kHasPrev | (finite ? kPrevFinite : 0u) | (cut ? 0u : kComputeRefl) | (sides << 4);
    ps.prevN = fr.z;
    ps.lcSum = lc.sum;
    ps.org = sp;
    ps.dir = ToWorld(fr, r.wil);
    ps.tmin = EpsRayCos(cosAbs);
    ps.len = len + 1;
    return true;
}

// One sample of PathTracer::EstimateIncomingRadiancePT with -sb 1, traced in one thread.
PG3_DEV F3 RadiancePT(const SceneView &S, const RngKey &key, F3 org, F3 dir, RayTally &t) {
    const LightSpec spec = SceneLightSpec(S);
    PathState ps;
    InitPath(ps, org, dir);
    for (;;) {
        Hit h;
        h.dist = 1e36f;
        const bool hit = Intersect(S, ps.org, ps.dir, ps.tmin, h);
        ++t.ext;
        ShadowReq sr;
        const bool cont = ShadeVertex(S, spec, key, ps, hit, h, sr, t.segments);
        if (sr.valid) {
            ++t.shadow;
            if (!Occluded(S, sr.p, sr.dir, sr.tmax)) ps.L += sr.contrib;
        }
        if (!cont) break;
    }
    return ps.L;
}

// PathTracer::EstimateIncomingRadiancePTNaive, path_tracer.hxx:45-165
PG3_DEV F3 RadiancePTNaive(const SceneView &S, const RngKey &key, F3 org, F3 dir, RayTally &t) {
    const LightSpec spec = SceneLightSpec(S);
    F3 out = mk(0.f), thr = mk(1.0f);
    float tmin = 0.f;
    for (uint32_t len = 1;; ++len) {
        Hit h;
        h.dist = 1e36f;
        ++t.ext;
        if (!Intersect(S, org, dir, tmin, h)) {
            if (len >= S.minPL && S.bgLight >= 0) out += BgRadiance(S, spec, dir) * thr;
            return out;
        }
        const F3 sp = org + dir * h.dist;
        const Frame fr = FrameFromZ(h.n);
        const F3 wol = ToLocal(fr, -dir);
        const DevMaterial mat = LoadGlobal(S.mats + h.mat);
        if (mat.lightId >= 0 && len >= S.minPL) {
            const DevLight Lt = LoadGlobal(S.lights + mat.lightId);
            if (Lt.type == kLightArea && wol.z > 0.f) out += Lt.radiance * thr;
        }
        if (ReflectanceZero(mat)) return out;
        if (S.maxPL > 0 && len >= S.maxPL) return out;
        if (len >= 1200u) return out;
        ++t.segments;
        VertexRng vr;
        RngVertex(key, len, vr);
        float rr = 1.0f;
        if (S.maxPL == 0) {
            rr = Clampf(RRProb(mat, wol), 0.0f, 0.997f);
            if (vr.u[kDimRR] > rr) return out;
        }
        BsdfRec r;
        SampleBsdf(mat, wol, vr.u[kDimBsdf0], vr.u[kDimBsdf1], vr.u[kDimBsdf2], r);
        if (IsBlocker(r)) return out;
        const float cosAbs = fabsf(r.wil.z);
        thr = thr * ((r.pdfW != kInf) ? (r.att * cosAbs) / (r.pdfW * rr * r.compProb) : r.att / (rr * r.compProb));
        org = sp;
        dir = ToWorld(fr, r.wil);
        tmin = EpsRayCos(cosAbs);
    }
}

// DirectIllumination::EstimateIncomingRadiance, direct_illumination.hxx:15-156
PG3_DEV F3 RadianceDirect(const SceneView &S, const RngKey &key, F3 org, F3 dir, RayTally &t) {
    const LightSpec spec = SceneLightSpec(S);
    Hit h;
    h.dist = 1e36f;
    ++t.ext;
    if (!Intersect(S, org, dir, 0.f, h)) return (S.bgLight >= 0) ? BgRadiance(S, spec, dir) : mk(0.f);
    ++t.segments;
    const F3 sp = org + dir * h.dist;
    const Frame fr = FrameFromZ(h.n);
    const F3 wol = ToLocal(fr, -dir);
    const DevMaterial mat = LoadGlobal(S.mats + h.mat);
    const uint32_t sides = mat.sides;
    const uint32_t vtx = 1;
    VertexRng vr;
    RngVertex(key, vtx, vr);
    LightCache lc;
    lc.sum = 0.f;
    if (S.numLights > 1 && S.algorithm != 3) FillLightCache(S, spec, sp, fr, sides, key, vtx, lc);
    F3 Lo = mk(0.f);
    const int alg = S.algorithm;
    if (alg == 3) {                                    // dlsa: every light once
        for (uint32_t i = 0; i < S.numLights; ++i) {
            LightSample ls;
            const float u0 = RngGet(key, vtx, i, kDimLight0), u1 = RngGet(key, vtx, i, kDimLight1),
                        u2 = RngGet(key, vtx, i, kDimLight2);
            if (SampleIllumination(S, spec, i, sp, fr, sides, u0, u1, u2, ls) && MaxC(ls.sample) > 0.f &&
                (++t.shadow, !Occluded(S, sp, ls.wig, ls.dist))) {                       // :315-348
                BsdfRec r;
                EvalBsdf(mat, ToLocal(fr, ls.wig), wol, r);
                Lo += (ls.pdfW != kInf) ? ls.sample * r.att / (ls.pdfW * 1.f) : ls.sample * r.att / 1.f;
            }
        }
    } else if (alg == 4 || alg == 5) {                 // dlss / dmis light part
        float prob = 1.f;
        const int32_t li = PickLight(spec, lc, vr.u[kDimLightPick], prob);
        LightSample ls;
        if (li >= 0 && SampleIllumination(S, spec, (uint32_t)li, sp, fr, sides, vr.u[kDimLight0], vr.u[kDimLight1],
                                          vr.u[kDimLight2], ls)) {
            ls.prob = prob;
            if (MaxC(ls.sample) > 0.f && (++t.shadow, !Occluded(S, sp, ls.wig, ls.dist))) {
                BsdfRec r;
                EvalBsdf(mat, ToLocal(fr, ls.wig), wol, r);
                if (alg == 4) {
                    Lo += (ls.pdfW != kInf) ? ls.sample * r.att / (ls.pdfW * ls.prob) : ls.sample * r.att / ls.prob;
                } else if (ls.pdfW != kInf) {
                    const float lightPdf = ls.pdfW * ls.prob;
                    Lo += (ls.sample * r.att * MisPower(lightPdf, r.pdfW * r.compProb)) / lightPdf;
                } else {
                    Lo += (ls.sample * r.att) / ls.prob;
                }
            }
        }
    }
    if (alg == 2 || alg == 5) {                        // dbs / dmis BSDF part
        BsdfRec r;
        SampleBsdf(mat, wol, vr.u[kDimBsdf0], vr.u[kDimBsdf1], vr.u[kDimBsdf2], r);
        if (!IsBlocker(r)) {
            // GetDirectRadianceFromDirection, path_tracer_base.hxx:92-154
            const F3 wig = ToWorld(fr, r.wil);
            Hit h2;
            h2.dist = 1e36f;
            F3 Li = mk(0.f);
            float lpdf = 0.f;
            int32_t lid = -1;
            ++t.ext;
            if (Intersect(S, sp, wig, EpsRayCos(r.wil.z), h2)) {
                const int32_t l2 = LoadGlobal(S.mats + h2.mat).lightId;
                if (l2 >= 0) {
                    const DevLight Lt = LoadGlobal(S.lights + l2);
                    const F3 lp = sp + wig * h2.dist;
                    const F3 lwol = ToLocal(FrameFromZ(h2.n), -wig);
                    if (Lt.type == kLightArea) {
                        if (lwol.z > 0.f) Li = Lt.radiance;
                        lpdf = AreaEmissionPdf(Lt, lp, sp);
                    }
                    lid = l2;
                }
            } else if (S.bgLight >= 0) {
                Li = BgRadiance(S, spec, wig);
                lpdf = BgPdf(S, spec, wig, fr, sides);
                lid = S.bgLight;
            }
            const float cosAbs = fabsf(r.wil.z);
            if (alg == 2) {
                Lo = (r.pdfW != kInf) ? (r.att * cosAbs * Li) / (r.pdfW * r.compProb) : (r.att * Li) / r.compProb;
            } else if (MaxC(Li) > 0.f) {                                                  // :427-490
                if (r.pdfW != kInf) {
                    const float bp = r.pdfW * r.compProb;
                    const float pick = (lid >= 0) ? PickProb(spec, lc, (uint32_t)lid) : 0.f;
                    Lo += (r.att * cosAbs * Li * MisPower(bp, lpdf * pick)) / bp;
                } else {
                    Lo += (r.att * Li) / (1.f * r.compProb);
                }
            }
        }
    }
    F3 Le = mk(0.f);
    if (mat.lightId >= 0) {
        const DevLight Lt = LoadGlobal(S.lights + mat.lightId);
        if (Lt.type == kLightArea && wol.z > 0.f) Le = Lt.radiance;
    }
    return Le + Lo;
}

// ---------------------------------------------------------------------------------------------
// EstimateIncomingRadiancePT with splitting (-sb > 1) and indirect clipping (-iic > 0): the
// recursion of path_tracer.hxx:167-430 run as a post-order walk over an explicit stack of
// vertices, since clipping needs each child's reflected estimate before it is weighted.
// ComputeSplittingCounts (:432-454) gives a vertex with budget b nB = round(1 - l + l b) BSDF
// samples and its children budget 1 + (b - nB) / nB = b / nB: nB never grows along a path, a
// vertex with nB >= 2 halves the budget at least, and a vertex with nB = 1 passes its budget on
// unchanged, so below at most 6 branching vertices (budgets <= 64) every subtree is a chain.  The
// walker keeps stack frames for the branching vertices only and evaluates a chain with
// ChainReflected, which folds a ring of per-vertex records and re-traces deep chains, so paths
// keep the reference's limit of 1200 vertices (kMaxPathLengthHard, :232-238).
constexpr uint32_t kTreeDepth = 8;            // stack frames: branching vertices (at most 6)
constexpr uint32_t kTreeChainRing = 32;       // per-thread fold records of a chain

// MISWeight2Power with sample counts, path_tracer_base.hxx:522-541
PG3_DEV float MisPowerN(float p1, uint32_t n1, float p2, uint32_t n2) {
    const float s1 = n1 * p1, s2 = n2 * p2;
    const float q1 = s1 * s1, q2 = s2 * s2;
    return FastDiv(q1, (q1 + q2) * n1);
}
// Vec3Base::ClipProportionally, math.hxx:538-549
PG3_DEV F3 ClipProportionally(F3 v, float maxAllowed) {
    const float m = MaxC(v);
    if (m > maxAllowed) {
        const float s = maxAllowed / m;
        v.x *= s;
        v.y *= s;
        v.z *= s;
    }
    return v;
}
// Indirect term of a BSDF sample: (att |cos| R) / (pdfW nB rr compProb), clipped proportionally
// to -iic (path_tracer.hxx:385-402).
PG3_DEV F3 ClippedIndirect(F3 attCos, F3 R, float den, float iic) {
    F3 ind = FastDiv(attCos * R, den);
    if (iic > 0.f) ind = ClipProportionally(ind, iic);
    return ind;
}

// Light samples l = l0, l0 + dl, ... < nLv of a shaded vertex (path_tracer.hxx:252-265,
// path_tracer_base.hxx:350-425) with MIS sample counts nLv, nBv; counts the shadow rays traced.
PG3_DEV F3 LightSampleSum(const SceneView &S, const LightSpec &spec, const RngKey &k, uint32_t vlen, F3 sp,
                          const Frame &fr, F3 wol, const DevMaterial &mat, uint32_t l0, uint32_t dl, uint32_t nLv,
                          uint32_t nBv, const LightCache &lc, uint32_t &shadows) {
    F3 acc = mk(0.f);
    if ((vlen + 1) < S.minPL) return acc;
#pragma unroll 1
    for (uint32_t l = l0; l < nLv; l += dl) {
        float prob = 1.f;
        const int32_t li = PickLight(spec, lc, RngGet(k, vlen, l, kDimLightPick), prob);
        LightSample ls;
        if (li >= 0 && SampleIllumination(S, spec, (uint32_t)li, sp, fr, mat.sides, RngGet(k, vlen, l, kDimLight0),
                                          RngGet(k, vlen, l, kDimLight1), RngGet(k, vlen, l, kDimLight2), ls)) {
            ls.prob = prob;
            if (MaxC(ls.sample) > 0.f && (++shadows, !Occluded(S, sp, ls.wig, ls.dist))) {
                BsdfRec r;
                EvalBsdf(mat, ToLocal(fr, ls.wig), wol, r);
                if (ls.pdfW != kInf) {
                    const float bsdfPdf = r.pdfW * r.compProb;
                    const float lightPdf = ls.pdfW * ls.prob;
                    acc += FastDiv(ls.sample * r.att * MisPowerN(lightPdf, nLv, bsdfPdf, nBv), lightPdf);
                } else {
                    acc += FastDiv(ls.sample * r.att, ls.prob * nLv);
                }
            }
        }
    }
    return acc;
}

// The parent vertex of an emission hit, as the MIS weight of that emission needs it
// (path_tracer.hxx:327-369): its position and normal, the factor att * |cos| (att for a Dirac
// sample) and pdfW * compProb (compProb for a Dirac sample) of its BSDF sample, its sample counts,
// light-cache sum, RNG key and vertex index.
struct EmissionParent {
    F3 sp, n, attCos;
    float pdf, lcSum;
    uint32_t flags;            // bit 0: finite BSDF sample; sides << 4
    uint32_t nB, nL, len;
    RngKey key;
};
// Emission `em` of light `lid` (pdf lpdf of the hit under light sampling) found by the parent's
// BSDF sample, MIS-weighted.
PG3_DEV F3 EmissionMis(const SceneView &S, const LightSpec &spec, F3 em, float lpdf, int32_t lid,
                       const EmissionParent &P) {
    if (IsZero(em) || lid < 0) return mk(0.f);
    if (!(P.flags & 1u)) return FastDiv(P.attCos * em, static_cast<float>(P.nB) * P.pdf);
    float pick = 1.f;
    if (spec.numLights == 0) {
        pick = 0.f;
    } else if (spec.numLights > 1) {
        pick = (P.lcSum > 0.f) ? EstimateContribution(S, spec, (uint32_t)lid, P.sp, FrameFromZ(P.n), P.flags >> 4, P.key,
                                                      P.len) / P.lcSum
                               : 1.f / spec.numLights;
    }
    return FastDiv(P.attCos * em * MisPowerN(P.pdf, P.nB, lpdf * pick, P.nL), P.pdf);
}

// Reflected radiance estimate of a shaded vertex whose subtree is a chain (every vertex one BSDF
// sample: path_tracer.hxx:240-404 with nB = 1): R_d = A_d + clip(W_d R_{d+1}), where A_d holds the
// vertex's light samples and the MIS-weighted emission its BSDF sample finds, and W_d that
// sample's weight.  The chain is traced forward pushing (A, att|cos|, denominator) per vertex on a
// ring of kTreeChainRing records and folded back to front at its end; when it is deeper than the
// ring, the folded estimate of the oldest record still held becomes the suffix of a re-trace from
// the chain's first vertex (the counter-based RNG replays it exactly), which stops at that depth:
// the same operations in the same order as one fold of the whole chain.  Vertex 0 of the chain is
// at `sp0` (normal n0, outgoing direction wol0, material mat0), is vertex `len0` of the path and
// uses RNG key k.
PG3_DEV F3 ChainReflected(const SceneView &S, const LightSpec &spec, const RngKey &k, uint32_t len0, F3 sp0, F3 n0,
                          F3 wol0, int32_t mat0, RayTally &t) {
    const uint32_t nL = (uint32_t)lroundf(fmaxf(S.lightRatio * 1u, 1.f));
    float ring[kTreeChainRing][7];      // depth d in slot d % kTreeChainRing: A.xyz, attCos.xyz, den
    int32_t target = -1;                // depth this pass stops at (-1: first pass)
    F3 rSuffix = mk(0.f);               // folded estimate of that depth
    for (;;) {
        F3 sp = sp0, wol = wol0, nrm = n0;
        int32_t mi = mat0;
        uint32_t d = 0;                 // depth of the vertex being shaded (vertex len0 + d)
        F3 tA = mk(0.f), tAttCos = mk(0.f), rChild = mk(0.f);
        float tDen = -1.f;
        for (;;) {                      // shade vertex d
            const uint32_t len = len0 + d;
            const Frame fr = FrameFromZ(nrm);
            const DevMaterial mat = LoadGlobal(S.mats + mi);
            if (target < 0) ++t.segments;
            LightCache lc;
            lc.sum = 0.f;
            if (spec.numLights > 1) FillLightCache(S, spec, sp, fr, mat.sides, k, len, lc);
            if (d > 0) {                // the parent's record is complete: push it
                float *r = ring[(d - 1) % kTreeChainRing];
                r[0] = tA.x;
                r[1] = tA.y;
                r[2] = tA.z;
                r[3] = tAttCos.x;
                r[4] = tAttCos.y;
                r[5] = tAttCos.z;
                r[6] = tDen;
            }
            tA = LightSampleSum(S, spec, k, len, sp, fr, wol, mat, 0u, 1u, nL, 1u, lc, t.shadow);
            const float rr = (S.maxPL == 0) ? Clampf(RRProb(mat, wol), 0.0f, 0.997f) : 1.0f;
            const bool cut = (S.maxPL == 0) && (RngGet(k, len, 0, kDimRR) > rr);
            BsdfRec r;
            SampleBsdf(mat, wol, RngGet(k, len, 0, kDimBsdf0), RngGet(k, len, 0, kDimBsdf1), RngGet(k, len, 0, kDimBsdf2),
                       r);
            if (IsBlocker(r)) {
                tDen = -1.f;
                break;
            }
            const bool finite = (r.pdfW != kInf);
            const float cosAbs = fabsf(r.wil.z);
            EmissionParent P;
            P.sp = sp;
            P.n = nrm;
            P.attCos = finite ? r.att * cosAbs : r.att;
            P.pdf = finite ? r.pdfW * r.compProb : r.compProb;
            P.lcSum = lc.sum;
            P.flags = (finite ? 1u : 0u) | (mat.sides << 4);
            P.nB = 1u;
            P.nL = nL;
            P.len = len;
            P.key = k;
            tAttCos = P.attCos;
            tDen = cut ? -1.f : (finite ? (r.pdfW * 1u * rr * r.compProb) : (1u * rr * r.compProb));
            // the child vertex d + 1 (:176-238, :408-422)
            const F3 dir = ToWorld(fr, r.wil);
            Hit h;
            h.dist = 1e36f;
            ++t.ext;
            const bool hit = Intersect(S, sp, dir, EpsRayCos(cosAbs), h);
            F3 em = mk(0.f);
            float lpdf = 0.f;
            int32_t lid = -1;
            const uint32_t clen = len + 1;
            DevMaterial cm{};
            F3 csp = mk(0.f), cwol = mk(0.f);
            if (hit) {
                csp = sp + dir * h.dist;
                cwol = ToLocal(FrameFromZ(h.n), -dir);
                cm = LoadGlobal(S.mats + h.mat);
                if (cm.lightId >= 0 && clen >= S.minPL) {
                    const DevLight Lt = LoadGlobal(S.lights + cm.lightId);
                    if (Lt.type == kLightArea) {
                        lpdf = AreaEmissionPdf(Lt, csp, sp);
                        if (cwol.z > 0.f) em = Lt.radiance;
                    } else if (Lt.type == kLightPoint) {
                        lpdf = kInf;
                    }
                    lid = cm.lightId;
                }
            } else if (clen >= S.minPL && S.bgLight >= 0) {
                em = BgRadiance(S, spec, dir);
                lpdf = BgPdf(S, spec, dir, fr, mat.sides);
                lid = S.bgLight;
            }
            tA += EmissionMis(S, spec, em, lpdf, lid, P);
            const bool atTarget = (target >= 0) && ((int32_t)d + 1 == target);
            if (atTarget) rChild = rSuffix;
            if (atTarget || !hit || cut || ReflectanceZero(cm) || (S.maxPL > 0 && clen >= S.maxPL) || clen >= 1200u)
                break;
            sp = csp;
            wol = cwol;
            nrm = h.n;
            mi = h.mat;
            ++d;
        }
        // fold: the newest record (depth d) is in registers, depths [0, d) in the ring
        F3 R = tA;
        if (tDen > 0.f && !IsZero(rChild)) R += ClippedIndirect(tAttCos, rChild, tDen, S.iic);
        const uint32_t lo = (d > kTreeChainRing) ? d - kTreeChainRing : 0u;
#pragma unroll 1
        for (uint32_t e = d; e > lo; --e) {
            const float *r = ring[(e - 1) % kTreeChainRing];
            F3 Rk = mk(r[0], r[1], r[2]);
            if (r[6] > 0.f && !IsZero(R)) Rk += ClippedIndirect(mk(r[3], r[4], r[5]), R, r[6], S.iic);
            R = Rk;
        }
        if (lo == 0u) return R;
        target = (int32_t)lo;           // depths [0, lo) were overwritten: re-trace up to depth lo
        rSuffix = R;
    }
}

struct TreeFrame {              // one vertex of the recursion, in its BSDF-sample loop
    F3 sp, n, wol;              // position, shading normal, outgoing direction (local)
    F3 refl;                    // oReflectedRadianceEstimate being accumulated
    F3 em;                      // emission found by the ray that reached this vertex (for the parent)
    F3 att;                     // record of the BSDF sample whose subtree is being traced
    float cosAbs, pdfW, comp;
    float nextBudget, rr, lcSum, lpdf;
    int32_t mat, lid;
    uint32_t len, branch, nB, nL, s, flags;   // flags: 1 cut (RR), 2 finite, sides << 4
};

PG3_DEV F3 RadiancePTTree(const SceneView &S, const LightSpec &spec, const RngKey &key0, F3 org, F3 dir,
                          RayTally &t) {
    TreeFrame stk[kTreeDepth];
    int32_t top = -1;
    // pending call EstimateIncomingRadiancePT(ray, len, computeRefl, budget) on branch `cBranch`
    F3 cOrg = org, cDir = dir;
    float cTmin = 0.f, cBudget = S.splitBudget;
    uint32_t cLen = 1, cBranch = 0;
    bool cRefl = true;
    // values returned by the last finished call
    F3 rEm = mk(0.f), rRefl = mk(0.f);
    float rPdf = 0.f;
    int32_t rLid = -1;
    enum { kCall, kNext, kReturn };
    int mode = kCall;
    for (;;) {
        if (mode == kCall) {
            RngKey key = key0;
            key.c3 = (key0.c3 & 0xFFFF0000u) | cBranch;
            rEm = mk(0.f);
            rRefl = mk(0.f);
            rPdf = 0.f;
            rLid = -1;
            mode = kReturn;
            Hit h;
            h.dist = 1e36f;
            ++t.ext;
            if (Intersect(S, cOrg, cDir, cTmin, h)) {
                const F3 sp = cOrg + cDir * h.dist;
                const Frame fr = FrameFromZ(h.n);
                const F3 wol = ToLocal(fr, -cDir);
                const DevMaterial mat = LoadGlobal(S.mats + h.mat);
                if (mat.lightId >= 0 && cLen >= S.minPL) {                            // :198-210
                    const DevLight Lt = LoadGlobal(S.lights + mat.lightId);
                    if (Lt.type == kLightArea) {
                        if (top >= 0) rPdf = AreaEmissionPdf(Lt, sp, cOrg);
                        if (wol.z > 0.f) rEm = Lt.radiance;
                    } else if (Lt.type == kLightPoint) {
                        rPdf = kInf;
                    }
                    rLid = mat.lightId;
                }
                // ComputeSplittingCounts, :432-454
                const uint32_t nB = (uint32_t)lroundf((1.f - S.splitLevel) * 1.f + S.splitLevel * cBudget);
                const uint32_t nL = (uint32_t)lroundf(fmaxf(S.lightRatio * nB, 1.f));
                // (a branching vertex beyond the stack cannot occur for budgets <= 64: cut defensively)
                const bool stop = !cRefl || ReflectanceZero(mat) || (S.maxPL > 0 && cLen >= S.maxPL) ||
                                  cLen >= 1200u || (nB != 1u && top + 1 >= (int32_t)kTreeDepth);
                if (!stop && nB == 1u) {
                    // one BSDF sample here, so in the whole subtree: a chain
                    rRefl = ChainReflected(S, spec, key, cLen, sp, fr.z, wol, h.mat, t);
                } else if (!stop) {
                    ++t.segments;
                    TreeFrame &F = stk[top + 1];
                    F.sp = sp;
                    F.n = fr.z;
                    F.wol = wol;
                    F.refl = mk(0.f);
                    F.em = rEm;
                    F.lpdf = rPdf;
                    F.lid = rLid;
                    F.mat = h.mat;
                    F.len = cLen;
                    F.branch = cBranch;
                    F.nB = nB;
                    F.nL = nL;
                    F.s = 0;
                    F.flags = mat.sides << 4;
                    F.nextBudget = 1 + ((cBudget - nB) / (float)nB);
                    LightCache lc;
                    lc.sum = 0.f;
                    if (spec.numLights > 1) FillLightCache(S, spec, sp, fr, mat.sides, key, cLen, lc);
                    F.lcSum = lc.sum;
                    if ((cLen + 1) >= S.minPL) {                                      // :252-265
                        for (uint32_t l = 0; l < nL; ++l) {
                            float prob = 1.f;
                            const int32_t li = PickLight(spec, lc, RngGet(key, cLen, l, kDimLightPick), prob);
                            LightSample ls;
                            if (li >= 0 &&
                                SampleIllumination(S, spec, (uint32_t)li, sp, fr, mat.sides, RngGet(key, cLen, l, kDimLight0),
                                                   RngGet(key, cLen, l, kDimLight1), RngGet(key, cLen, l, kDimLight2), ls)) {
                                ls.prob = prob;                                       // :350-425
                                if (MaxC(ls.sample) > 0.f && (++t.shadow, !Occluded(S, sp, ls.wig, ls.dist))) {
                                    BsdfRec r;
                                    EvalBsdf(mat, ToLocal(fr, ls.wig), wol, r);
                                    if (ls.pdfW != kInf) {
                                        const float bsdfPdf = r.pdfW * r.compProb;
                                        const float lightPdf = ls.pdfW * ls.prob;
                                        F.refl += FastDiv(ls.sample * r.att * MisPowerN(lightPdf, nL, bsdfPdf, nB), lightPdf);
                                    } else {
                                        F.refl += FastDiv(ls.sample * r.att, ls.prob * nL);
                                    }
                                }
                            }
                        }
                    }
                    F.rr = (S.maxPL == 0) ? Clampf(RRProb(mat, wol), 0.0f, 0.997f) : 1.0f;
                    ++top;
                    mode = kNext;
                }
            } else if (cLen >= S.minPL && S.bgLight >= 0) {                           // :408-422
                rEm = BgRadiance(S, spec, cDir);
                if (top >= 0) rPdf = BgPdf(S, spec, cDir, FrameFromZ(stk[top].n), stk[top].flags >> 4);
                rLid = S.bgLight;
            }
        } else if (mode == kNext) {                                                   // :281-404
            TreeFrame &F = stk[top];
            mode = kReturn;
            RngKey key = key0;
            key.c3 = (key0.c3 & 0xFFFF0000u) | F.branch;
            const DevMaterial mat = LoadGlobal(S.mats + F.mat);
            while (F.s < F.nB) {
                const uint32_t s = F.s;
                const bool cut = (S.maxPL == 0) && (RngGet(key, F.len, s, kDimRR) > F.rr);
                BsdfRec r;
                SampleBsdf(mat, F.wol, RngGet(key, F.len, s, kDimBsdf0), RngGet(key, F.len, s, kDimBsdf1),
                           RngGet(key, F.len, s, kDimBsdf2), r);
                if (IsBlocker(r)) {
                    ++F.s;
                    continue;
                }
                const bool finite = (r.pdfW != kInf);
                F.att = r.att;
                F.cosAbs = fabsf(r.wil.z);
                F.pdfW = r.pdfW;
                F.comp = r.compProb;
                F.flags = (F.flags & ~3u) | (cut ? 1u : 0u) | (finite ? 2u : 0u);
                cOrg = F.sp;
                cDir = ToWorld(FrameFromZ(F.n), r.wil);
                cTmin = EpsRayCos(F.cosAbs);
                cLen = F.len + 1;
                cRefl = !cut;
                cBudget = F.nextBudget;
                cBranch = (F.nB > 1) ? ((F.branch * 7u + s + 1u) & 0xFFFFu) : F.branch;
                mode = kCall;
                break;
            }
            if (mode == kReturn) {            // every BSDF sample done: this vertex returns
                rEm = F.em;
                rRefl = F.refl;
                rPdf = F.lpdf;
                rLid = F.lid;
                --top;
            }
        } else {                                                                      // kReturn
            if (top < 0) return rEm + rRefl;
            TreeFrame &P = stk[top];
            const bool finite = P.flags & 2u, cut = P.flags & 1u;
            if (!IsZero(rEm) && rLid >= 0) {
                if (finite) {
                    float pick = 1.f;
                    if (spec.numLights == 0) {
                        pick = 0.f;
                    } else if (spec.numLights > 1) {
                        if (P.lcSum > 0.f) {
                            RngKey key = key0;
                            key.c3 = (key0.c3 & 0xFFFF0000u) | P.branch;
                            pick = EstimateContribution(S, spec, (uint32_t)rLid, P.sp, FrameFromZ(P.n), P.flags >> 4,
                                                        key, P.len) / P.lcSum;
                        } else {
                            pick = 1.f / spec.numLights;
                        }
                    }
                    const float lp = rPdf * pick;
                    const float bp = P.pdfW * P.comp;
                    P.refl += FastDiv(P.att * P.cosAbs * rEm * MisPowerN(bp, P.nB, lp, P.nL), bp);
                } else {
                    P.refl += FastDiv(P.att * rEm, static_cast<float>(P.nB) * P.comp);
                }
            }
            if (!cut && !IsZero(rRefl)) {
                F3 ind;
                if (finite) ind = FastDiv(P.att * P.cosAbs * rRefl, P.pdfW * P.nB * P.rr * P.comp);
                else ind = FastDiv(P.att * rRefl, P.nB * P.rr * P.comp);
                if (S.iic > 0.f) ind = ClipProportionally(ind, S.iic);
                P.refl += ind;
            }
            ++P.s;
            mode = kNext;
        }
    }
}

// Camera sample `iteration` of `pixel`: jitter, primary ray (PathTracerBase::RunIteration,
// path_tracer_base.hxx:66-80) and the RNG key of its path.
PG3_DEV void CameraSample(const SceneView &S, uint32_t pixel, uint32_t iteration, RngKey &key, F3 &org, F3 &dir) {
    key = RngKey{S.k0, S.k1, pixel, iteration, S.sceneId << 16};
    const float jx = RngGet(key, 0, 0, 0), jy = RngGet(key, 0, 0, 1);   // path_tracer_base.hxx:74-77
    const uint32_t x = pixel % S.resX, y = pixel / S.resX;
    GenerateRay(S, float(x) + jx, float(y) + jy, org, dir);
}

// The same for camera `view` of a multi-view scene (null: the scene's own camera).
PG3_DEV void CameraSample(const SceneView &S, const DevView *view, uint32_t pixel, uint32_t iteration, RngKey &key,
                          F3 &org, F3 &dir) {
    if (!view) {
        CameraSample(S, pixel, iteration, key, org, dir);
        return;
    }
    key = RngKey{S.k0, S.k1, pixel, iteration, view->sceneId << 16};
    const float jx = RngGet(key, 0, 0, 0), jy = RngGet(key, 0, 0, 1);
    const uint32_t x = pixel % S.resX, y = pixel / S.resX;
    GenerateRay(view->r2w, view->camPos, float(x) + jx, float(y) + jy, org, dir);
}

// Stages the inner levels of a steerable EM's triangle tree in `lds` (12 floats per node as F4v
// triples, at most kMaxNodes nodes: the levels from the root down that fit, which are
// contiguous at the end of nodeW), cooperatively by the workgroup; the caller synchronises the
// workgroup before the returned cache is read.
constexpr uint32_t kTreeLdsNodes = 1280;
// TreeLevels: which levels StageTree copies (minLevel == E.numLevels: none fits) and how many nodes.
template <uint32_t kMaxNodes = kTreeLdsNodes>
PG3_DEV uint32_t TreeLevels(const DevEm &E, uint32_t &count) {
    uint32_t minLevel = E.numLevels;
    count = 0;
    for (int32_t l = (int32_t)E.numLevels - 1; l >= 1; --l) {
        if (count + E.levelCount[l] > kMaxNodes) break;
        count += E.levelCount[l];
        minLevel = (uint32_t)l;
    }
    return minLevel;
}
// The copy is shared by the lanes first, first + stride, ... (the workgroup by default).
template <uint32_t kMaxNodes = kTreeLdsNodes>
PG3_DEV TreeCache StageTree(const DevEm &E, F4v *lds, uint32_t first = threadIdx.x, uint32_t stride = blockDim.x) {
    TreeCache tc;
    uint32_t count;
    const uint32_t minLevel = TreeLevels<kMaxNodes>(E, count);
    if (minLevel < E.numLevels) {
        const uint32_t base = E.levelBase[minLevel];
        const GlobalPtr<F4v> src = (GlobalPtr<F4v>)AsGlobal(E.nodeW + (size_t)base * 12u);
        for (uint32_t i = first; i < count * 3u; i += stride) lds[i] = src[i];
        tc.nodes = reinterpret_cast<const float *>(lds);
        tc.minLevel = minLevel;
        tc.baseNode = base;
    }
    return tc;
}

// Stages the simple sampler's marginal CDF (h + 1 floats, searched by every light sample) in
// `lds` when it fits in kMargLdsMax entries (an EM of up to 4095 rows), cooperatively by the
// workgroup; the caller synchronises the workgroup before the staged table is read.
constexpr uint32_t kMargLdsMax = 4096;
PG3_DEV void StageMarginal(const DevEm &E, float *lds, TreeCache &tc, uint32_t first = threadIdx.x,
                           uint32_t stride = blockDim.x) {
    if (E.h + 1u > kMargLdsMax) return;
    const GlobalPtr<float> src = AsGlobal(E.margCdf);
    for (uint32_t i = first; i <= E.h; i += stride) lds[i] = src[i];
    tc.marg = lds;
}

// A work item of PathTracer -sb 1: the 8x8 pixel tile with corner (x0, y0) of scene `scene`
// x iterations [it, it + total / 64).  Its samples q = 0 .. total-1 are (tile pixel q % 64,
// iteration it + q / 64).
struct TileItem {
    uint32_t scene, x0, y0, it, total;
    double *fb;
};

// Persistent wavefront loop over work items.  Tasks are handed out in order, one to each lane
// whose task has ended (path regeneration across pixels and iterations), and a lane whose task
// ends adds its radiance to its pixel of the framebuffer at once (FP64 atomics), so an item is
// done with as soon as it has handed out its last task: no lane waits for another lane's path,
// however long (Russian roulette, deep chains).  The lanes of a wavefront share one scene (its
// tables are read through scalar loads): with kMultiScene, an item of another scene waits until
// the running lanes have finished.  All bookkeeping is wave-uniform.  Which lane traces which
// task changes nothing: a sample's path depends only on its (pixel, iteration) RNG stream.
//   src.Next(TileItem&) -> bool   fetches the next item (false: none left; wave-uniform)
//   src.View(scene) -> const SceneView *, src.Spec(scene) -> LightSpec
//   kMultiScene: src.Enter(scene) when the wavefront starts items of a scene (all lanes idle),
//   src.Entered(scene) -> whether it did for the current one
//   Path: a lane's task.  Path::Subs(S) tasks per camera sample (sub-samples whose radiances add
//   up to the sample's); Start(S, pixel, iteration, sub); Step(S, spec, tally, shadowed, step, ring)
//   -> true when the task is done, with `step` the wavefront's step number (the same on every lane)
//   and `ring` the lane's Path::Ring, an array kept apart from the path so that the path's own
//   fields stay in registers (an array indexed at run time lives in scratch memory, and a struct
//   holding one would be kept there whole);
//   Result() -> its radiance; CountsSample() -> whether the task counts as a camera sample.  Step
//   traces exactly one extension ray (counted here) and counts its segments; a path with at most
//   one shadow ray per step sets `shadowed` when it traced one (counted here), others count
//   their own in the tally.
// FP64 adds to global memory and to LDS.  The explicit address space keeps them global and LDS
// atomics where the compiler cannot tell where a pointer points (a framebuffer pointer read from
// memory, an LDS array passed through a struct): a flat atomic may address either, so every
// later wait for an LDS access would also wait for it.
PG3_DEV void GlobalAdd(double *p, double v) {
    __hip_atomic_fetch_add((__attribute__((address_space(1))) double *)p, v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
PG3_DEV void LdsAdd(double *p, double v) {
    __hip_atomic_fetch_add((__attribute__((address_space(3))) double *)p, v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class Path, bool kMultiScene, class Src>
PG3_DEV void TraceTasks(Src &src, RayTally &tally) {
    const uint32_t lane = __lane_id();
    TileItem T{};                      // item handing out tasks (wave-uniform)
    uint32_t next = 0;                 // its first task not handed out
    uint32_t scene = 0;                // scene of the running lanes
    bool more = true;
    bool busy = false;
    uint32_t px = 0;                   // this lane's pixel
    double *fb = nullptr;              // and its framebuffer
    uint32_t step = 0;                 // steps taken by the wavefront (wave-uniform)
    Path path;
    typename Path::Ring ring;
    for (;;) {
        const uint64_t idle = __ballot(!busy);
        if (idle != 0 && more && next >= T.total) {
            if (src.Next(T)) {
                T.total *= Path::Subs(*src.View(T.scene));
                next = 0;
            } else {
                more = false;
            }
        }
        if constexpr (kMultiScene) {   // drained: switch
            if (idle == ~0ull && (T.scene != scene || !src.Entered(T.scene))) {
                scene = T.scene;
                src.Enter(scene);
            }
            src.Join();                // a cached copy of the scene's tables, once one is staged
        }
        // the scene is looked up by index at each use: a pointer variable carried around the
        // loop would lose its address space (flat loads in place of scalar and LDS loads)
        const SceneView &S = *src.View(scene);
        const LightSpec spec = src.Spec(scene);
        if (idle != 0 && next < T.total && (!kMultiScene || T.scene == scene)) {
            const uint32_t subs = Path::Subs(S);
            const uint32_t q = next + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
            if (!busy && q < T.total) {
                const uint32_t pix = q & 63u;
                const uint32_t x = T.x0 + (pix & 7u), y = T.y0 + (pix >> 3);
                if (x < S.resX && y < S.resY) {
                    const uint32_t r = q >> 6;
                    px = y * S.resX + x;
                    fb = T.fb;
                    path.Start(S, px, T.it + r / subs, r % subs);
                    busy = true;
                }
            }
            next += (uint32_t)__popcll(idle);
        }
        const uint64_t stepping = __ballot(busy);
        if (stepping == 0) {
            if (!more && next >= T.total) break;
            continue;
        }
        tally.wExt += (uint32_t)__popcll(stepping);        // one Intersect per Step
        bool shadowed = false, counted = false;
        if (busy && path.Step(S, spec, tally, shadowed, step, ring)) {
            const F3 L = path.Result();
            double *f = fb + (size_t)px * 3;
            GlobalAdd(f + 0, (double)L.x);
            GlobalAdd(f + 1, (double)L.y);
            GlobalAdd(f + 2, (double)L.z);
            busy = false;
            counted = path.CountsSample();
        }
        tally.wSamples += (uint32_t)__popcll(__ballot(counted));
        tally.wShadow += (uint32_t)__popcll(__ballot(shadowed));
        ++step;
    }
}

// Per-lane fields of a task.  kLds: in LDS, one slot per lane of a workgroup of at most
// kLaneFieldLanes lanes (field i of lane x at i * kLaneFieldLanes + x: consecutive lanes,
// consecutive banks); else in registers.
constexpr uint32_t kLaneFieldLanes = 1024;
template <uint32_t N, bool kLds>
struct LaneFields {
    uint32_t v[N];
    PG3_DEV uint32_t Get(uint32_t i) const { return v[i]; }
    PG3_DEV void Set(uint32_t i, uint32_t x) { v[i] = x; }
};
template <uint32_t N>
struct LaneFields<N, true> {
    PG3_DEV static uint32_t *Base() {
        __shared__ uint32_t fields[N * kLaneFieldLanes];
        return fields;
    }
    PG3_DEV uint32_t Get(uint32_t i) const { return Base()[i * kLaneFieldLanes + threadIdx.x]; }
    PG3_DEV void Set(uint32_t i, uint32_t x) { Base()[i * kLaneFieldLanes + threadIdx.x] = x; }
};

// A lane's sample of the path tracer with -sb 1 -iic 0: the chain of ShadeVertex, one vertex
// (extension ray, light sample and its shadow ray, BSDF sample) per step.
// kLds: the sample's pixel and iteration and its radiance so far live in LDS (LaneFields); a step
// rebuilds the RNG key from them and the scene, and adds the radiance it finds to the LDS sum.
template <bool kMergeEm, bool kLds = false>
struct ChainPath {
    struct Keyed {                      // kLds = false: the RNG key in registers
        RngKey key;
    };
    struct Unkeyed {};
    using KeyStore = typename std::conditional<kLds, Unkeyed, Keyed>::type;
    enum : uint32_t { kPixel, kIteration, kL, kFields = kL + 3 };
    KeyStore ks;
    LaneFields<kLds ? kFields : 1u, kLds> cold;
    PathState ps;
    struct Ring {};
    PG3_DEV static constexpr uint32_t Subs(const SceneView &) { return 1u; }
    PG3_DEV static constexpr bool CountsSample() { return true; }
    PG3_DEV void Start(const SceneView &S, uint32_t pixel, uint32_t iteration, uint32_t) {
        F3 org, dir;
        RngKey key;
        CameraSample(S, pixel, iteration, key, org, dir);
        InitPath(ps, org, dir);
        if constexpr (kLds) {
            cold.Set(kPixel, pixel);
            cold.Set(kIteration, iteration);
            cold.Set(kL, 0u);
            cold.Set(kL + 1, 0u);
            cold.Set(kL + 2, 0u);
        } else {
            ks.key = key;
        }
    }
    PG3_DEV bool Step(const SceneView &S, const LightSpec &spec, RayTally &t, bool &shadowed, uint32_t, Ring &) {
        RngKey key;
        if constexpr (kLds) {
            key = RngKey{S.k0, S.k1, cold.Get(kPixel), cold.Get(kIteration), S.sceneId << 16};
            ps.L = mk(0.f);
        } else {
            key = ks.key;
        }
        Hit h;
        h.dist = 1e36f;
        const bool hit = Intersect(S, ps.org, ps.dir, ps.tmin, h);
        ShadowReq sr;
        const bool cont = ShadeVertex<kMergeEm>(S, spec, key, ps, hit, h, sr, t.segments);
        shadowed = sr.valid;
        if (sr.valid && !Occluded(S, sr.p, sr.dir, sr.tmax)) ps.L += sr.contrib;
        if constexpr (kLds) {
            if (!IsZero(ps.L)) {
                cold.Set(kL, __float_as_uint(__uint_as_float(cold.Get(kL)) + ps.L.x));
                cold.Set(kL + 1, __float_as_uint(__uint_as_float(cold.Get(kL + 1)) + ps.L.y));
                cold.Set(kL + 2, __float_as_uint(__uint_as_float(cold.Get(kL + 2)) + ps.L.z));
            }
        }
        return !cont;
    }
    PG3_DEV F3 Result() const {
        if constexpr (kLds)
            return mk(__uint_as_float(cold.Get(kL)), __uint_as_float(cold.Get(kL + 1)), __uint_as_float(cold.Get(kL + 2)));
        else
            return ps.L;
    }
};

// A lane's share of one sample of EstimateIncomingRadiancePT (path_tracer.hxx:167-430) with
// splitting at the first vertex only and indirect clipping: the case of every integer splitting
// budget at the reference's splitting level 1 (its defaults -sb 4 -iic 16), where
// ComputeSplittingCounts (:432-454) gives the first vertex nB BSDF and nL light samples and
// budget 1 to every later vertex.  The first vertex's estimate is its light samples plus, per BSDF
// sample s, the MIS-weighted emission s finds and the clipped indirect estimate of s's subtree, a
// chain; these terms add, so a sample splits into nB sub-samples that run on their own lanes:
// sub-sample s traces the camera ray, takes the first vertex's light samples l = s, s + nB, ...
// and its BSDF sample s, and follows s's chain.  Every lane runs the same per-vertex code (one
// light-sample set and one BSDF sample per step), and the pixel's sum adds the sub-samples.
// Clipping makes a vertex's estimate R_d = A_d + clip(W_d R_{d+1}) depend on the whole suffix (A_d:
// its light samples and the emission its BSDF sample finds, W_d: that sample's weight), so each
// vertex's record is pushed on a per-lane ring and the chain's end folds the ring back to front.
// The ring slot of a push is the wavefront's step number (a lane pushes at most once per step), so
// the lanes of a wavefront store to and load from the same slot: scratch accesses coalesce into
// whole cache lines.  A chain deeper than the ring (Russian roulette lets a path bounce inside a
// dielectric for many vertices; the reference stops at 1200) folds the records the ring still
// holds, re-traces its sub-sample from the camera ray up to the folded depth (the counter-based RNG
// replays it exactly) and folds the next part onto the folded suffix, and so on to the first
// vertex: the same operations in the same order as one fold of the whole chain.  RNG streams as
// RadiancePTTree (the chain of BSDF sample s is branch s + 1 when nB > 1).
// kMergeEm: the scene's only light is a steerable EM and every lane takes at most one light
// sample per step (LightCount(nB) <= nB at the first vertex and 1 at the others: the reference's
// light-to-BSDF ratio 1).  The escaped lanes' emission and the shading lanes' light samples then
// share one pass of EM lookups (SteerableEmQuery), and the light sample's shadow ray is traced at
// the end of the step, after the BSDF sample, when little else is live; otherwise the light
// samples of a vertex run in a loop with their shadow rays (LightSampleSum).
#ifndef PG3_CHAIN_RING
#define PG3_CHAIN_RING 32
#endif
constexpr uint32_t kChainRing = PG3_CHAIN_RING;
#ifndef PG3_FOLD_BATCH
#define PG3_FOLD_BATCH 4
#endif
constexpr uint32_t kFoldBatch = PG3_FOLD_BATCH;   // ring records loaded together by a fold
static_assert(kChainRing >= 2 && (kChainRing & (kChainRing - 1)) == 0,
              "ring slots follow the step counter modulo a power of two, across its wrap-around");
template <bool kMergeEm>
struct SplitPath {
    // ray in flight: it reaches vertex `len` (1: the camera ray's hit) of sub-sample `sub`
    F3 org, dir;
    float tmin;
    uint32_t len;
    bool refl;                   // the vertex it reaches is shaded (not cut by Russian roulette)
    // newest record (vertex len - 1): A, the indirect factors of its BSDF sample (attCos, den;
    // den < 0: no indirect term -- cut by Russian roulette or no sample), and what the MIS weight
    // of the emission found at vertex len needs of that vertex (its position is `org`): its normal,
    // pdfW * compProb of its BSDF sample (compProb for a Dirac sample), light-cache sum, flags
    F3 tA, tAttCos, pN;
    float tDen, pPdf, pLcSum;
    uint32_t pFlags;             // bit 0: finite BSDF sample; sides << 4
    // Fields a step reads a few times at most: the sample's RNG key (pixel, iteration), the
    // sub-sample index, the count nPush of records of depths [0, nPush) (vertex = depth + 1) pushed
    // at consecutive steps, the depth `target` a re-trace of a deep chain stops at (-1: first pass)
    // and that depth's estimate rSuffix, and the sub-sample's radiance L (sub-sample 0 adds the
    // camera hit's emission).  The split kernel of a steerable EM keeps them in LDS, which frees
    // the registers it would otherwise spill.
    enum : uint32_t { kPixel, kIteration, kSub, kPush, kTarget, kSuffix, kL = kSuffix + 3, kFields = kL + 3 };
    LaneFields<kFields, kMergeEm> cold;
    PG3_DEV uint32_t pixel() const { return cold.Get(kPixel); }
    PG3_DEV uint32_t iteration() const { return cold.Get(kIteration); }
    PG3_DEV uint32_t sub() const { return cold.Get(kSub); }
    PG3_DEV uint32_t nPush() const { return cold.Get(kPush); }
    PG3_DEV int32_t target() const { return (int32_t)cold.Get(kTarget); }
    PG3_DEV F3 Vec(uint32_t f) const {
        return mk(__uint_as_float(cold.Get(f)), __uint_as_float(cold.Get(f + 1)), __uint_as_float(cold.Get(f + 2)));
    }
    PG3_DEV void SetVec(uint32_t f, F3 v) {
        cold.Set(f, __float_as_uint(v.x));
        cold.Set(f + 1, __float_as_uint(v.y));
        cold.Set(f + 2, __float_as_uint(v.z));
    }
    // A record: the vertex's A and the indirect weight W = attCos / den of its BSDF sample (0
    // without an indirect term), as one 16-byte and one 8-byte access.
    struct Ring {
        float4 a[kChainRing];   // A.xyz, W.x
        float2 b[kChainRing];   // W.yz
    };
    PG3_DEV static F3 IndirectWeight(F3 attCos, float den) { return den > 0.f ? FastDiv(attCos, den) : mk(0.f); }
    // R_d = A_d + the indirect estimate W_d R_{d+1}, clipped (path_tracer.hxx:380-404)
    PG3_DEV static F3 FoldRecord(F3 A, F3 W, F3 R, float iic) {
        if (IsZero(W) || IsZero(R)) return A;
        F3 ind = W * R;
        if (iic > 0.f) ind = ClipProportionally(ind, iic);
        return A + ind;
    }

    // BSDF samples of the first vertex (ComputeSplittingCounts, :440-446): the sub-samples per sample
    PG3_DEV static uint32_t Subs(const SceneView &S) {
        return (uint32_t)lroundf((1.f - S.splitLevel) * 1.f + S.splitLevel * S.splitBudget);
    }
    PG3_DEV static uint32_t LightCount(const SceneView &S, uint32_t nBv) {
        return (uint32_t)lroundf(fmaxf(S.lightRatio * nBv, 1.f));
    }
    PG3_DEV bool CountsSample() const { return sub() == 0u; }
    PG3_DEV void Start(const SceneView &S, uint32_t px, uint32_t it, uint32_t s) {
        RngKey key;
        CameraSample(S, px, it, key, org, dir);
        cold.Set(kPixel, px);
        cold.Set(kIteration, it);
        cold.Set(kSub, s);
        tmin = 0.f;
        len = 1;
        refl = true;
        SetVec(kL, mk(0.f));
        cold.Set(kPush, 0u);
        cold.Set(kTarget, (uint32_t)-1);
    }
    PG3_DEV RngKey Branch(const SceneView &S, uint32_t b) const {
        return RngKey{S.k0, S.k1, pixel(), iteration(), (S.sceneId << 16) | b};
    }
    // RNG branch of vertex v of sub-sample s
    PG3_DEV static uint32_t BranchOf(uint32_t v, uint32_t nB, uint32_t s) {
        return (v == 1u || nB <= 1u) ? 0u : ((s + 1u) & 0xFFFFu);
    }
    // MIS-weighted emission that the BSDF sample of vertex len - 1 finds at vertex len
    PG3_DEV F3 ParentEmission(const SceneView &S, const LightSpec &spec, F3 em, float lpdf, int32_t lid,
                              uint32_t nB, uint32_t sb) const {
        EmissionParent P;
        P.sp = org;
        P.n = pN;
        P.attCos = tAttCos;
        P.pdf = pPdf;
        P.lcSum = pLcSum;
        P.flags = pFlags;
        P.nB = (len == 2u) ? nB : 1u;
        P.nL = LightCount(S, P.nB);
        P.len = len - 1u;
        P.key = Branch(S, BranchOf(len - 1u, nB, sb));
        return EmissionMis(S, spec, em, lpdf, lid, P);
    }
    // End of a chain pass at wavefront step t.  The newest record is in registers and rChild is the
    // estimate of the vertex after it (0, or the folded suffix at a re-trace's target); the ring
    // holds the older records, the newest of them pushed at step t - jStart.  Folds them back to
    // front; if the oldest records were overwritten, sets up the re-trace up to the oldest folded
    // depth and returns false, else adds the first vertex's estimate to L and returns true.
    PG3_DEV bool Fold(const SceneView &S, uint32_t t, uint32_t jStart, F3 rChild, const Ring &ring) {
        F3 R = FoldRecord(tA, IndirectWeight(tAttCos, tDen), rChild, S.iic);
        const uint32_t np = nPush(), n = min(np, kChainRing);
        // kFoldBatch records per memory round trip: their loads are issued together (a slot past
        // the chain's oldest record is still inside the ring and its value is not used)
#pragma unroll 1
        for (uint32_t i = 0; i < n; i += kFoldBatch) {
            float4 a[kFoldBatch];
            float2 b[kFoldBatch];
#pragma unroll
            for (uint32_t j = 0; j < kFoldBatch; ++j) {
                const uint32_t slot = (t - jStart - i - j) % kChainRing;
                a[j] = ring.a[slot];
                b[j] = ring.b[slot];
            }
#pragma unroll
            for (uint32_t j = 0; j < kFoldBatch; ++j)
                if (i + j < n) R = FoldRecord(mk(a[j].x, a[j].y, a[j].z), mk(a[j].w, b[j].x, b[j].y), R, S.iic);
        }
        if (np > kChainRing) {
            cold.Set(kTarget, np - kChainRing);
            SetVec(kSuffix, R);
            RngKey key;
            CameraSample(S, pixel(), iteration(), key, org, dir);
            tmin = 0.f;
            len = 1;
            refl = true;
            cold.Set(kPush, 0u);
            return false;
        }
        SetVec(kL, Vec(kL) + R);
        return true;
    }
    // One vertex per step (t: the wavefront's step number, the same on every lane).
    PG3_DEV bool Step(const SceneView &S, const LightSpec &spec, RayTally &tally, bool &shadowed, uint32_t t,
                      Ring &ring) {
        if constexpr (kMergeEm) return StepMerged(S, spec, tally, shadowed, t, ring);
        else return StepGeneric(S, spec, tally, t, ring);
    }
    // Ring record of the parent vertex (its A is complete once the emission its BSDF sample finds
    // is added), pushed at step t.
    PG3_DEV void Push(uint32_t t, Ring &ring) {
        const F3 W = IndirectWeight(tAttCos, tDen);
        ring.a[t % kChainRing] = make_float4(tA.x, tA.y, tA.z, W.x);
        ring.b[t % kChainRing] = make_float2(W.y, W.z);
        cold.Set(kPush, nPush() + 1u);
    }
    // BSDF sample si of vertex len (:281-404) with nBv samples at the vertex: the indirect factors
    // of the new record, what the MIS weight of the emission it finds needs, and the next ray.
    // Returns false for a blocker sample (no indirect term).
    PG3_DEV bool SampleNext(const SceneView &S, const RngKey &k, uint32_t si, uint32_t nBv, F3 sp, const Frame &fr,
                            F3 wol, const DevMaterial &mat) {
        const float rr = (S.maxPL == 0) ? Clampf(RRProb(mat, wol), 0.0f, 0.997f) : 1.0f;
        const bool cut = (S.maxPL == 0) && (RngGet(k, len, si, kDimRR) > rr);
        BsdfRec r;
        SampleBsdf(mat, wol, RngGet(k, len, si, kDimBsdf0), RngGet(k, len, si, kDimBsdf1), RngGet(k, len, si, kDimBsdf2),
                   r);
        if (IsBlocker(r)) {
            tDen = -1.f;
            return false;
        }
        const bool finite = (r.pdfW != kInf);
        const float cosAbs = fabsf(r.wil.z);
        tAttCos = finite ? r.att * cosAbs : r.att;
        tDen = cut ? -1.f : (finite ? (r.pdfW * nBv * rr * r.compProb) : (nBv * rr * r.compProb));
        pPdf = finite ? r.pdfW * r.compProb : r.compProb;
        pN = fr.z;
        pFlags = (finite ? 1u : 0u) | (mat.sides << 4);
        org = sp;
        dir = ToWorld(fr, r.wil);
        tmin = EpsRayCos(cosAbs);
        len = len + 1;
        refl = !cut;
        return true;
    }
    PG3_DEV bool StepMerged(const SceneView &S, const LightSpec &spec, RayTally &tally, bool &shadowed, uint32_t t,
                            Ring &ring) {
        const uint32_t nB = Subs(S), sb = sub();
        const int32_t tgt = target();
        const RngKey k = Branch(S, BranchOf(len, nB, sb));
        Hit h;
        h.dist = 1e36f;
        const bool hit = Intersect(S, org, dir, tmin, h);
        const bool first = (len == 1);
        const bool atTarget = (tgt >= 0) && ((int32_t)len - 1 == tgt);
        F3 em = mk(0.f);                 // emission found at vertex len (area light or EM)
        float lpdf = 0.f;
        int32_t lid = -1;
        F3 sp = mk(0.f), wol = mk(0.f);
        Frame fr{};
        DevMaterial mat{};
        bool shade = false, emq = false, nee = false;
        F3 qn = pN;
        uint32_t qsides = pFlags >> 4;
        float uSide = 0.f, u0 = 0.f, u1 = 0.f, prob = 1.f;
        const uint32_t nBv = first ? nB : 1u, nLv = LightCount(S, nBv), l = first ? sb : 0u;
        if (hit) {                                                                    // :176-238
            sp = org + dir * h.dist;
            fr = FrameFromZ(h.n);
            wol = ToLocal(fr, -dir);
            mat = LoadGlobal(S.mats + h.mat);
            if (mat.lightId >= 0 && len >= S.minPL) {
                const DevLight Lt = LoadGlobal(S.lights + mat.lightId);
                if (Lt.type == kLightArea) {
                    if (len > 1) lpdf = AreaEmissionPdf(Lt, sp, org);
                    if (wol.z > 0.f) em = Lt.radiance;
                } else if (Lt.type == kLightPoint) {
                    lpdf = kInf;
                }
                lid = mat.lightId;
            }
            shade = !atTarget && refl && !ReflectanceZero(mat) && !(S.maxPL > 0 && len >= S.maxPL) && len < 1200u;
            if (shade && (len + 1) >= S.minPL && l < nLv) {                          // :252-265
                LightCache lc;
                lc.sum = 0.f;
                const int32_t li = PickLight(spec, lc, RngGet(k, len, l, kDimLightPick), prob);
                if (li >= 0) {
                    emq = nee = true;
                    qn = fr.z;
                    qsides = mat.sides;
                    u0 = RngGet(k, len, l, kDimLight0);
                    u1 = RngGet(k, len, l, kDimLight1);
                    uSide = RngGet(k, len, l, kDimLight2);
                }
            }
        } else {
            emq = len >= S.minPL && S.bgLight >= 0;                                   // :408-422
        }
        EmQuery q;
        q.Le = mk(0.f);
        q.emPdf = 0.f;
        q.lsOk = false;
        if (emq) SteerableEmQuery(S.em, spec, nee, qn, dir, qsides, uSide, u0, u1, prob, q);
        if (!hit && emq) {
            em = q.Le;
            lpdf = q.emPdf;
            lid = S.bgLight;
        }
        if (first) {
            if (sb == 0u && tgt < 0) SetVec(kL, Vec(kL) + em);   // the camera hit's emission, once per sample
        } else {
            tA += ParentEmission(S, spec, em, lpdf, lid, nB, sb);
        }
        if (!shade) {
            if (first) return true;
            return Fold(S, t, 1u, atTarget ? Vec(kSuffix) : mk(0.f), ring);
        }
        if (tgt < 0 && (!first || sb == 0u)) ++tally.segments;   // the shared first vertex once
        if (!first) Push(t, ring);                      // the parent's record is complete
        // the light sample's contribution (path_tracer_base.hxx:350-425); its shadow ray is traced
        // after the BSDF sample
        F3 contrib = mk(0.f), sdir = mk(0.f);
        float smax = 0.f;
        bool sv = false;
        if (q.lsOk && MaxC(q.ls.sample) > 0.f) {
            BsdfRec r;
            EvalBsdf(mat, ToLocal(fr, q.ls.wig), wol, r);
            if (q.ls.pdfW != kInf) {
                const float bsdfPdf = r.pdfW * r.compProb;
                const float lightPdf = q.ls.pdfW * q.ls.prob;
                contrib = FastDiv(q.ls.sample * r.att * MisPowerN(lightPdf, nLv, bsdfPdf, nBv), lightPdf);
            } else {
                contrib = FastDiv(q.ls.sample * r.att, q.ls.prob * nLv);
            }
            sdir = q.ls.wig;
            smax = q.ls.dist;
            sv = true;
        }
        tA = mk(0.f);
        pLcSum = 0.f;
        const bool more = SampleNext(S, k, first ? sb : 0u, nBv, sp, fr, wol, mat);
        shadowed = sv;
        if (sv && !Occluded(S, sp, sdir, smax)) tA = contrib;
        if (!more) return Fold(S, t, 0u, mk(0.f), ring);
        return false;
    }
    PG3_DEV bool StepGeneric(const SceneView &S, const LightSpec &spec, RayTally &tally, uint32_t t, Ring &ring) {
        const uint32_t nB = Subs(S), sb = sub();
        const int32_t tgt = target();
        const RngKey k = Branch(S, BranchOf(len, nB, sb));
        Hit h;
        h.dist = 1e36f;
        const bool hit = Intersect(S, org, dir, tmin, h);
        F3 em = mk(0.f);
        float lpdf = 0.f;
        int32_t lid = -1;
        F3 sp = mk(0.f), wol = mk(0.f);
        Frame fr{};
        DevMaterial mat{};
        if (hit) {                                                                    // :176-238
            sp = org + dir * h.dist;
            fr = FrameFromZ(h.n);
            wol = ToLocal(fr, -dir);
            mat = LoadGlobal(S.mats + h.mat);
            if (mat.lightId >= 0 && len >= S.minPL) {
                const DevLight Lt = LoadGlobal(S.lights + mat.lightId);
                if (Lt.type == kLightArea) {
                    if (len > 1) lpdf = AreaEmissionPdf(Lt, sp, org);
                    if (wol.z > 0.f) em = Lt.radiance;
                } else if (Lt.type == kLightPoint) {
                    lpdf = kInf;
                }
                lid = mat.lightId;
            }
        } else if (len >= S.minPL && S.bgLight >= 0) {                               // :408-422
            em = BgRadiance(S, spec, dir);
            if (len > 1) lpdf = BgPdf(S, spec, dir, FrameFromZ(pN), pFlags >> 4);
            lid = S.bgLight;
        }
        if (len == 1) {
            if (sb == 0u && tgt < 0) SetVec(kL, Vec(kL) + em);   // the camera hit's emission, once per sample
        } else {
            tA += ParentEmission(S, spec, em, lpdf, lid, nB, sb);
        }
        const bool atTarget = (tgt >= 0) && ((int32_t)len - 1 == tgt);
        const bool stop = atTarget || !hit || !refl || ReflectanceZero(mat) || (S.maxPL > 0 && len >= S.maxPL) ||
                          len >= 1200u;
        if (stop) {
            if (len == 1) return true;
            return Fold(S, t, 1u, atTarget ? Vec(kSuffix) : mk(0.f), ring);
        }
        const bool first = (len == 1);
        if (tgt < 0 && (!first || sb == 0u)) ++tally.segments;   // the shared first vertex once
        if (!first) Push(t, ring);                      // the parent's record is complete
        LightCache lc;
        lc.sum = 0.f;
        if (spec.numLights > 1) FillLightCache(S, spec, sp, fr, mat.sides, k, len, lc);
        const uint32_t nBv = first ? nB : 1u;
        tA = LightSampleSum(S, spec, k, len, sp, fr, wol, mat, first ? sb : 0u, nBv, LightCount(S, nBv), nBv, lc,
                            tally.shadow);
        pLcSum = lc.sum;
        if (!SampleNext(S, k, first ? sb : 0u, nBv, sp, fr, wol, mat)) return Fold(S, t, 0u, mk(0.f), ring);
        return false;
    }
    PG3_DEV F3 Result() const { return Vec(kL); }
};

// Item source of one scene: items chunk * numTiles + tile from the launch's work counter.
template <bool kMerge>
struct SceneItems {
    static constexpr bool kMergeEm = kMerge;
    const SceneView *S;
    LightSpec spec;
    uint32_t *work;
    double *fb;
    uint32_t firstIter, nIter, itersPerItem, numItems, numTiles, tilesX;
    // items [0, mainItems) cover iterations [0, mainIters) in chunks of itersPerItem; the
    // remaining iterations follow as single-iteration items, so the items handed out last are
    // the shortest and the waves' finishing times even out at the end of the launch
    uint32_t mainItems, mainIters;
    PG3_DEV bool Next(TileItem &t) {
        uint32_t item = 0;
        if (__lane_id() == 0) item = atomicAdd(work, 1u);
        item = __builtin_amdgcn_readfirstlane(__shfl(item, 0, 64));
        if (item >= numItems) return false;   // every wave leaves once the counter passes the last item
        uint32_t tile, n;
        if (item < mainItems) {
            tile = item % numTiles;
            t.it = firstIter + (item / numTiles) * itersPerItem;
            n = min(firstIter + nIter, t.it + itersPerItem) - t.it;
        } else {
            const uint32_t j = item - mainItems;
            tile = j % numTiles;
            t.it = firstIter + mainIters + j / numTiles;
            n = 1u;
        }
        t.scene = 0;
        t.x0 = (tile % tilesX) * 8u;
        t.y0 = (tile / tilesX) * 8u;
        t.total = 64u * n;
        t.fb = fb;
        return true;
    }
    // Splits the last chunk (when items hold more than one iteration) into single iterations.
    PG3_DEV void SetTail(uint32_t numChunks) {
        const uint32_t mainChunks = (itersPerItem > 1u && numChunks > 1u) ? numChunks - 1u : numChunks;
        mainIters = min(nIter, mainChunks * itersPerItem);
        mainItems = numTiles * mainChunks;
        numItems = mainItems + numTiles * (nIter - mainIters);
    }
    PG3_DEV const SceneView *View(uint32_t) const { return S; }
    PG3_DEV LightSpec Spec(uint32_t) const { return spec; }
};

PG3_DEV F3 RadianceSample(const SceneView &S, uint32_t pixel, uint32_t iteration, RayTally &t) {
    RngKey key;
    F3 org, dir;
    CameraSample(S, pixel, iteration, key, org, dir);
    if (S.algorithm == 7) return RadiancePT(S, key, org, dir, t);
    if (S.algorithm == 6) return RadiancePTNaive(S, key, org, dir, t);
    return RadianceDirect(S, key, org, dir, t);
}

}  // namespace pg3
