// Per-sample radiance estimators of the reference renderers, written as iterative device code.
// PathTracer::EstimateIncomingRadiancePT (path_tracer.hxx:167-430) with splitting budget 1 is a
// chain: pixel = emitted(1) + sum_v B_v * (NEE_v + E_v), where B_v is the path throughput up to
// vertex v, NEE_v the MIS-weighted light sample (path_tracer_base.hxx:350-425) and E_v the
// MIS-weighted emission found by the BSDF sample of v (path_tracer.hxx:327-369).  The sums are
// accumulated forward; the indirect clipping of path_tracer.hxx:400-402 needs the whole suffix and
// is therefore rejected by the host when -iic > 0.
#pragma once

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

// Vertex of the path tracer with a steerable EM; the same results as the generic vertex below,
// in another order.  An escaped path (emission with MIS, path_tracer.hxx:408-422: the EM pdf of
// its direction at the previous vertex) and a light sample of the EM (SampleImpl and
// SampleHemisphereImpl, em_steerable_sampler.hxx:2084-2217) need the same lookups: the
// hemisphere's clamped-cosine coefficients and integrals, the EM radiance at one direction and
// the luminance at its antipode (one lat-long mapping), and the two GetBasePdf terms
// (:2387-2413) of that direction and its antipode.  Lanes of both kinds run these lookups in one
// pass, so a wavefront no longer issues them once for its escaped lanes and again for its
// shading lanes.
PG3_DEV bool ShadeVertexSteerable(const SceneView &S, const LightSpec &spec, const RngKey &key, PathState &ps,
                                  bool hit, const Hit &h, ShadowReq &sr, uint32_t &segments) {
    const DevEm &E = S.em;
    const uint32_t len = ps.len;
    F3 sp = mk(0.f), wol = mk(0.f);
    Frame fr{};
    DevMaterial mat{};
    VertexRng vr{};
    LightCache lc;
    lc.sum = 0.f;
    bool shade = false;            // hit lane that shades this vertex
    bool emq = false, nee = false;  // lane runs the EM lookups; they serve a light sample
    F3 qn = ps.prevN, qdir = ps.dir;
    uint32_t qsides = ps.flags >> 4;
    float uSide = 0.f, u0 = 0.f, u1 = 0.f, prob = 1.f;
    if (hit) {
        sp = ps.org + ps.dir * h.dist;
        fr = FrameFromZ(h.n);
        wol = ToLocal(fr, -ps.dir);
        mat = LoadGlobal(S.mats + h.mat);
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
    F3 Le = mk(0.f);
    float emPdf = 0.f;   // escaped lane: pdf of its direction under the previous vertex's EM sampling
    LightSample ls;
    bool lsOk = false;
    if (emq) {
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
        Le = EmEvalUV(E, u, v);
        Antipode(u, v, uo, vo);
        const float lumOpp = EmLumUV(E, uo, vo);
        // GetBasePdf of qdir and of -qdir; a light sample needs the one of -(sampled direction)
        const float t1 = BasePdfLum(E, qdir, Luminance(Le), c);
        const float t2 = BasePdfLum(E, -qdir, lumOpp, c);
        if (nee) {
            lsOk = ok;
            float pdf = triPdf;
            pdf += flip ? t1 : t2;
            if (front && back) pdf *= FastDiv(up ? ui : li, wi);
            ls.wig = qdir;
            ls.pdfW = pdf;
            ls.sample = Le * cs;
            ls.prob = prob;
            ls.dist = 3.402823466e+38f;   // em_sampler.hxx:57
        } else if (ok) {
            const float p = (Dot(qdir, hn) < 0.f) ? 0.f : t1 + t2;
            emPdf = (front && back) ? p * FastDiv(up ? ui : li, wi) : p;
        }
    }
    if (!hit) {                                                                       // :408-422
        if (emq) {
            if (!(ps.flags & kHasPrev)) {
                ps.L += Le;
            } else if (!IsZero(Le)) {
                if (ps.flags & kPrevFinite) {
                    const float lp = emPdf * PrevPickProb(S, spec, key, ps, (uint32_t)S.bgLight);
                    ps.L += ps.eThr * Le * MisPower(ps.prevBsdfPdf, lp);
                } else {
                    ps.L += ps.eThr * Le;
                }
            }
        }
        return false;
    }
    if (!shade) return false;
    if (lsOk) LightSampleRequest(ps, mat, fr, wol, sp, ls, sr);
    return ContinuePath(S, ps, sp, fr, wol, mat, vr, lc.sum);
}

// One vertex of EstimateIncomingRadiancePT (path_tracer.hxx:167-430) with one light sample and
// one BSDF sample: adds the MIS-weighted emission found by the previous BSDF sample, emits the
// light-sample request and samples the BSDF.  Returns true if the path continues.
// kMergeEm: kernels whose only light is a steerable EM take ShadeVertexSteerable (with several
// lights its extra live state costs more register spill than it saves).
template <bool kMergeEm = false>
PG3_DEV bool ShadeVertex(const SceneView &S, const LightSpec &spec, const RngKey &key, PathState &ps, bool hit, const Hit &h,
                         ShadowReq &sr, uint32_t &segments) {
    sr.valid = false;
    if constexpr (kMergeEm) return ShadeVertexSteerable(S, spec, key, ps, hit, h, sr, segments);
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
    const DevMaterial mat = LoadGlobal(S.mats + h.mat);
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
// vertices, since clipping needs each child's reflected estimate before it is weighted.  Paths
// are cut at kTreeDepth vertices (the reference cuts at 1200, kMaxPathLengthHard); with
// -maxpl <= kTreeDepth the two agree exactly.  (SplitPath, the first-vertex splitting of the
// reference's defaults, re-traces deep chains instead and keeps the reference's 1200.)
constexpr uint32_t kTreeDepth = 48;

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
                const bool stop = !cRefl || ReflectanceZero(mat) || (S.maxPL > 0 && cLen >= S.maxPL) ||
                                  cLen >= 1200u || top + 1 >= (int32_t)kTreeDepth;
                if (!stop) {
                    ++t.segments;
                    TreeFrame &F = stk[top + 1];
                    // ComputeSplittingCounts, :432-454
                    const uint32_t nB = (uint32_t)lroundf((1.f - S.splitLevel) * 1.f + S.splitLevel * cBudget);
                    const uint32_t nL = (uint32_t)lroundf(fmaxf(S.lightRatio * nB, 1.f));
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

// Stages the inner levels of a steerable EM's triangle tree in `lds` (12 floats per node as F4v
// triples, at most kMaxNodes nodes: the levels from the root down that fit, which are
// contiguous at the end of nodeW), cooperatively by the workgroup; the caller synchronises the
// workgroup before the returned cache is read.
constexpr uint32_t kTreeLdsNodes = 1280;
template <uint32_t kMaxNodes = kTreeLdsNodes>
PG3_DEV TreeCache StageTree(const DevEm &E, F4v *lds) {
    TreeCache tc;
    uint32_t minLevel = E.numLevels, count = 0;
    for (int32_t l = (int32_t)E.numLevels - 1; l >= 1; --l) {
        if (count + E.levelCount[l] > kMaxNodes) break;
        count += E.levelCount[l];
        minLevel = (uint32_t)l;
    }
    if (minLevel < E.numLevels) {
        const uint32_t base = E.levelBase[minLevel];
        const GlobalPtr<F4v> src = (GlobalPtr<F4v>)AsGlobal(E.nodeW + (size_t)base * 12u);
        for (uint32_t i = threadIdx.x; i < count * 3u; i += blockDim.x) lds[i] = src[i];
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
PG3_DEV void StageMarginal(const DevEm &E, float *lds, TreeCache &tc) {
    if (E.h + 1u > kMargLdsMax) return;
    const GlobalPtr<float> src = AsGlobal(E.margCdf);
    for (uint32_t i = threadIdx.x; i <= E.h; i += blockDim.x) lds[i] = src[i];
    tc.marg = lds;
}

// Adds LDS slot `slot` (64 pixels x 3 sums of one tile) to the framebuffer and clears it.
PG3_DEV void FlushSlot(double *slot, uint32_t x0, uint32_t y0, uint32_t resX, uint32_t resY, double *fb) {
    const uint32_t lane = __lane_id();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t x = x0 + (lane & 7u), y = y0 + (lane >> 3);
    if (x < resX && y < resY) {
        double *p = fb + (size_t)(y * resX + x) * 3;
        atomicAdd(p + 0, slot[lane * 3 + 0]);
        atomicAdd(p + 1, slot[lane * 3 + 1]);
        atomicAdd(p + 2, slot[lane * 3 + 2]);
    }
    slot[lane * 3 + 0] = 0.0;
    slot[lane * 3 + 1] = 0.0;
    slot[lane * 3 + 2] = 0.0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A work item of PathTracer -sb 1: the 8x8 pixel tile with corner (x0, y0) of scene `scene`
// x iterations [it, it + total / 64).  Its samples q = 0 .. total-1 are (tile pixel q % 64,
// iteration it + q / 64).
struct TileItem {
    uint32_t scene, x0, y0, it, total;
    double *fb;
};

// Persistent wavefront loop over work items.  Samples are handed out in order, one to each lane
// whose path has ended (path regeneration across pixels and iterations), and the wavefront keeps
// two items open: once the newer item has handed out its last sample, lanes whose paths end
// start samples of a further item in the other slot while the older item's last paths still
// run, so no lane waits for the longest path of an item.  Each item's sums are gathered in its
// LDS slot (`acc`: 2 x 64 pixels x 3 FP64) and added to its framebuffer as soon as its last
// sample ends.  Open items share one scene: an item of another scene waits until the older item
// has drained.  All slot bookkeeping is wave-uniform.  Which lane traces which sample changes
// nothing: a sample's path depends only on its (pixel, iteration) RNG stream.
//   src.Next(TileItem&) -> bool   fetches the next item (false: none left; wave-uniform)
//   src.View(scene) -> const SceneView *, src.Spec(scene) -> LightSpec
//   Path: a lane's sample; Start(S, pixel, iteration), Step(S, spec, tally, shadowed) -> true
//   when the sample is done, Result() -> its radiance.  Step traces exactly one extension ray
//   (counted here) and counts its segments; a path with at most one shadow ray per step sets
//   `shadowed` when it traced one (counted here), others count their own in the tally.
template <class Path, class Src>
PG3_DEV void TraceItems(Src &src, double *acc, RayTally &tally) {
    const uint32_t lane = __lane_id();
#pragma unroll
    for (uint32_t k = 0; k < 6; ++k) acc[lane + 64u * k] = 0.0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    TileItem A{}, B{};                 // the two slots (wave-uniform)
    uint32_t penA = 0, penB = 0;       // samples of the slot in flight
    bool liveA = false, liveB = false, more = true;
    uint32_t cur = 0, next = 0;        // slot handing out samples, its first sample not handed out
    bool held = false;                 // cur's scene differs from the scene of the running paths
    uint32_t scene = 0;                // scene of the running paths
    bool busy = false;
    uint32_t pix = 0, slot = 0;        // this lane's sample: tile pixel and slot
    Path path;
    for (;;) {
        const uint64_t idle = __ballot(!busy);
        if (idle != 0 && more && !(liveA && liveB) && !((cur ? liveB : liveA) && next < (cur ? B.total : A.total))) {
            TileItem t;
            if (!src.Next(t)) more = false;
            else {
                cur = liveA ? 1u : 0u;
                if (cur) B = t;
                else A = t;
                next = 0;
                held = (t.scene != scene) && (liveA || liveB);
                if (cur) { penB = 0; liveB = true; }
                else { penA = 0; liveA = true; }
            }
        }
        if (held && !(cur ? liveA : liveB)) held = false;   // the older item has drained
        const bool curOpen = !held && (cur ? liveB : liveA) && next < (cur ? B.total : A.total);
        if (curOpen) scene = cur ? B.scene : A.scene;
        // the scene is looked up by index at each use: a pointer variable carried around the
        // loop would lose its address space (flat loads in place of scalar and LDS loads)
        const SceneView &S = *src.View(scene);
        const LightSpec spec = src.Spec(scene);
        if (idle != 0 && curOpen) {
            const TileItem &T = cur ? B : A;
            const uint32_t q = next + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
            bool start = false;
            if (!busy && q < T.total) {
                pix = q & 63u;
                const uint32_t x = T.x0 + (pix & 7u), y = T.y0 + (pix >> 3);
                if (x < S.resX && y < S.resY) {
                    path.Start(S, y * S.resX + x, T.it + (q >> 6));
                    busy = true;
                    slot = cur;
                    start = true;
                }
            }
            next += (uint32_t)__popcll(idle);
            const uint32_t started = (uint32_t)__popcll(__ballot(start));
            if (cur) penB += started;
            else penA += started;
        }
        const uint64_t stepping = __ballot(busy);
        if (stepping != 0) {
            tally.wExt += (uint32_t)__popcll(stepping);    // one Intersect per Step
            bool done = false, shadowed = false;
            if (busy) {
                if (path.Step(S, spec, tally, shadowed)) {
                    const F3 L = path.Result();
                    double *a = acc + slot * 192u + pix * 3u;
                    atomicAdd(a + 0, (double)L.x);
                    atomicAdd(a + 1, (double)L.y);
                    atomicAdd(a + 2, (double)L.z);
                    busy = false;
                    done = true;
                }
            }
            const uint64_t fin = __ballot(done), finB = __ballot(done && slot != 0);
            tally.wSamples += (uint32_t)__popcll(fin);
            tally.wShadow += (uint32_t)__popcll(__ballot(shadowed));
            penB -= (uint32_t)__popcll(finB);
            penA -= (uint32_t)__popcll(fin & ~finB);
        }
        // an item is complete once it has handed out every sample and none is in flight (a held
        // item has handed out none)
        if (liveA && penA == 0 && !(cur == 0 && (held || next < A.total))) {
            const SceneView *V = src.View(A.scene);
            FlushSlot(acc, A.x0, A.y0, V->resX, V->resY, A.fb);
            liveA = false;
        }
        if (liveB && penB == 0 && !(cur == 1 && (held || next < B.total))) {
            const SceneView *V = src.View(B.scene);
            FlushSlot(acc + 192, B.x0, B.y0, V->resX, V->resY, B.fb);
            liveB = false;
        }
        if (!more && !liveA && !liveB) break;
    }
}

// A lane's sample of the path tracer with -sb 1 -iic 0: the chain of ShadeVertex, one vertex
// (extension ray, light sample and its shadow ray, BSDF sample) per step.
template <bool kMergeEm>
struct ChainPath {
    RngKey key;
    PathState ps;
    PG3_DEV void Start(const SceneView &S, uint32_t pixel, uint32_t iteration) {
        F3 org, dir;
        CameraSample(S, pixel, iteration, key, org, dir);
        InitPath(ps, org, dir);
    }
    PG3_DEV bool Step(const SceneView &S, const LightSpec &spec, RayTally &t, bool &shadowed) {
        Hit h;
        h.dist = 1e36f;
        const bool hit = Intersect(S, ps.org, ps.dir, ps.tmin, h);
        ShadowReq sr;
        const bool cont = ShadeVertex<kMergeEm>(S, spec, key, ps, hit, h, sr, t.segments);
        shadowed = sr.valid;
        if (sr.valid && !Occluded(S, sr.p, sr.dir, sr.tmax)) ps.L += sr.contrib;
        return !cont;
    }
    PG3_DEV F3 Result() const { return ps.L; }
};

// A lane's sample of EstimateIncomingRadiancePT (path_tracer.hxx:167-430) with splitting at the
// first vertex only and indirect clipping: the case of every integer splitting budget at the
// reference's splitting level 1 (its defaults -sb 4 -iic 16), where ComputeSplittingCounts
// (:432-454) gives the first vertex nB BSDF and nL light samples and budget 1 to every later
// vertex.  The first vertex's estimate is its light samples plus, per BSDF sample, the MIS-
// weighted emission found and the clipped indirect estimate of that sample's subtree; the terms
// of different BSDF samples add, so the nB subtrees are traced one after another, each a chain.
// Clipping makes a chain vertex's estimate R_v = A_v + clip(w_v R_{v+1}) depend on the whole
// suffix (A_v: its light samples and the emission its BSDF sample finds), so each vertex pushes
// (A_v, the factors of w_v) on a per-lane ring of kChainStack records and the chain's end folds
// them back to front.  A chain deeper than the ring (Russian roulette lets a path bounce inside
// a dielectric for many vertices; the reference stops at 1200) folds its deepest kChainStack
// records, then re-traces the chain from its start up to the folded depth (the counter-based RNG
// replays it exactly) and folds the next segment onto the folded suffix, and so on to the root:
// the same operations in the same order as one fold of the whole chain.  One vertex per step,
// the same RNG streams, operations and order as RadiancePTTree, which handles every other
// splitting configuration.
constexpr uint32_t kChainStack = kTreeDepth - 1;
struct SplitPath {
    // ray in flight: it reaches vertex `len` on RNG branch `branch`
    F3 org, dir;
    float tmin;
    uint32_t len, branch;
    bool refl;                   // the vertex it reaches is shaded (not cut by Russian roulette)
    RngKey key;                  // branch 0 (c3 low half = 0)
    // first vertex
    F3 rSp, rN, rWol, em1, reflRoot;
    int32_t rMat;
    float rRR, rLcSum;
    uint32_t nB, nL, s, rSides;
    F3 rAttCos;                  // current first-vertex BSDF sample: clip((rAttCos * R) / rDen)
    float rDen;
    bool rCut;
    // parent of the vertex the ray reaches (MIS of the emission found there)
    F3 pAttCos, pSp, pN;
    float pPdfW, pComp, pLcSum;
    uint32_t pSides, pLen, pBranch, pnB, pnL;
    bool pFinite;
    // chain ring: record d (chain depth d = vertex d + 2) in slot d % kChainStack; `top` is the
    // newest record's depth (-1: none) and its A is kept in topA.  Re-tracing a deep chain:
    // `target` is the depth at which the pass stops (-1: first pass) and rSuffix the folded
    // estimate of the chain beyond it.
    int32_t top, target;
    uint64_t cutMask;            // per slot: the record's BSDF sample was cut by Russian roulette
    F3 topA, rSuffix;
    float stA[kChainStack][3], stW[kChainStack][3], stDen[kChainStack];
    F3 L;

    PG3_DEV void Start(const SceneView &S, uint32_t pixel, uint32_t iteration) {
        CameraSample(S, pixel, iteration, key, org, dir);
        tmin = 0.f;
        len = 1;
        branch = 0;
        refl = true;
        em1 = mk(0.f);
        reflRoot = mk(0.f);
        top = -1;
        target = -1;
    }
    PG3_DEV static uint32_t Slot(int32_t d) { return (uint32_t)d % kChainStack; }
    PG3_DEV RngKey Branch(uint32_t b) const {
        RngKey k = key;
        k.c3 = (key.c3 & 0xFFFF0000u) | b;
        return k;
    }
    // Light-sample estimate of a shaded vertex (:252-265, :350-425) with nLv light and nBv BSDF
    // samples (MIS counts).
    PG3_DEV F3 LightSamples(const SceneView &S, const LightSpec &spec, const RngKey &k, uint32_t vlen, F3 sp,
                            const Frame &fr, F3 wol, const DevMaterial &mat, uint32_t nLv, uint32_t nBv,
                            const LightCache &lc, uint32_t &shadows) const {
        F3 acc = mk(0.f);
        if ((vlen + 1) < S.minPL) return acc;
        for (uint32_t l = 0; l < nLv; ++l) {
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
    // BSDF sample `sub` of a vertex; false for a blocker.  Sets the ray to the next vertex, the
    // parent record for that vertex's emission, and the indirect-term factors (attCos, den, cut).
    PG3_DEV bool SampleAt(const SceneView &S, const RngKey &k, uint32_t vlen, uint32_t vbranch, uint32_t sub, F3 sp,
                          const Frame &fr, F3 wol, const DevMaterial &mat, uint32_t nBv, uint32_t nLv, float rr,
                          float lcSum, F3 &attCos, float &den, bool &cut) {
        cut = (S.maxPL == 0) && (RngGet(k, vlen, sub, kDimRR) > rr);
        BsdfRec r;
        SampleBsdf(mat, wol, RngGet(k, vlen, sub, kDimBsdf0), RngGet(k, vlen, sub, kDimBsdf1),
                   RngGet(k, vlen, sub, kDimBsdf2), r);
        if (IsBlocker(r)) return false;
        const bool finite = (r.pdfW != kInf);
        const float cosAbs = fabsf(r.wil.z);
        attCos = finite ? r.att * cosAbs : r.att;
        den = finite ? (r.pdfW * nBv * rr * r.compProb) : (nBv * rr * r.compProb);
        pAttCos = attCos;
        pPdfW = r.pdfW;
        pComp = r.compProb;
        pFinite = finite;
        pSp = sp;
        pN = fr.z;
        pSides = mat.sides;
        pLen = vlen;
        pBranch = vbranch;
        pnB = nBv;
        pnL = nLv;
        pLcSum = lcSum;
        org = sp;
        dir = ToWorld(fr, r.wil);
        tmin = EpsRayCos(cosAbs);
        len = vlen + 1;
        branch = (nBv > 1) ? ((vbranch * 7u + sub + 1u) & 0xFFFFu) : vbranch;
        refl = !cut;
        return true;
    }
    // MIS-weighted emission delivered to the parent (path_tracer.hxx:327-369)
    PG3_DEV F3 ParentEmission(const SceneView &S, const LightSpec &spec, F3 em, float lpdf, int32_t lid) const {
        if (IsZero(em) || lid < 0) return mk(0.f);
        if (!pFinite) return FastDiv(pAttCos * em, static_cast<float>(pnB) * pComp);
        float pick = 1.f;
        if (spec.numLights == 0) {
            pick = 0.f;
        } else if (spec.numLights > 1) {
            pick = (pLcSum > 0.f) ? EstimateContribution(S, spec, (uint32_t)lid, pSp, FrameFromZ(pN), pSides,
                                                         Branch(pBranch), pLen) / pLcSum
                                  : 1.f / spec.numLights;
        }
        const float lp = lpdf * pick;
        const float bp = pPdfW * pComp;
        return FastDiv(pAttCos * em * MisPowerN(bp, pnB, lp, pnL), bp);
    }
    PG3_DEV static F3 Indirect(F3 attCos, F3 R, float den, float iic) {
        F3 ind = FastDiv(attCos * R, den);
        if (iic > 0.f) ind = ClipProportionally(ind, iic);
        return ind;
    }
    // End of a chain pass (records [0, top] pushed): fold the records the ring still holds onto
    // the suffix beyond them.  If older records were overwritten, set up the re-trace up to the
    // oldest folded depth and return false (the caller restarts the chain from the same
    // first-vertex sample); otherwise add the clipped indirect term of the subtree and return true.
    PG3_DEV bool FoldChain(const SceneView &S) {
        const int32_t lo = max(0, top + 1 - (int32_t)kChainStack);
        F3 R = (target >= 0) ? rSuffix : mk(0.f);
        for (int32_t k = top; k >= lo; --k) {
            const uint32_t sl = Slot(k);
            const F3 A = (k == top) ? topA : mk(stA[sl][0], stA[sl][1], stA[sl][2]);
            F3 Rk = A;
            if (!((cutMask >> sl) & 1ull) && !IsZero(R))
                Rk += Indirect(mk(stW[sl][0], stW[sl][1], stW[sl][2]), R, stDen[sl], S.iic);
            R = Rk;
        }
        top = -1;
        if (lo > 0) {
            rSuffix = R;
            target = lo;
            return false;
        }
        if (!rCut && !IsZero(R)) reflRoot += Indirect(rAttCos, R, rDen, S.iic);
        target = -1;
        ++s;
        return true;
    }
    PG3_DEV bool Step(const SceneView &S, const LightSpec &spec, RayTally &t, bool &) {
        const RngKey k = Branch(branch);
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
            if (len > 1) lpdf = BgPdf(S, spec, dir, FrameFromZ(pN), pSides);
            lid = S.bgLight;
        }
        // (a re-traced pass delivers the emission of vertex 2 to the first vertex only once)
        if (len == 1) em1 = em;
        else if (len == 2) { if (target < 0) reflRoot += ParentEmission(S, spec, em, lpdf, lid); }
        else topA += ParentEmission(S, spec, em, lpdf, lid);
        const bool atTarget = (target >= 0) && ((int32_t)len - 2 == target);
        const bool stop = atTarget || !hit || !refl || ReflectanceZero(mat) || (S.maxPL > 0 && len >= S.maxPL) ||
                          len >= 1200u;
        // what the step samples next: the first vertex's next BSDF sample or the chain vertex's one
        bool wantRoot = false, wantChain = false;
        float rr = 1.0f;
        LightCache lc;
        lc.sum = 0.f;
        if (stop) {
            if (len == 1) {
                L = em1 + reflRoot;
                return true;
            }
            FoldChain(S);
            wantRoot = true;
        } else {
            if (target < 0) ++t.segments;
            if (spec.numLights > 1) FillLightCache(S, spec, sp, fr, mat.sides, k, len, lc);
            const bool first = (len == 1);
            if (first) {                                                              // :240-250
                nB = (uint32_t)lroundf((1.f - S.splitLevel) * 1.f + S.splitLevel * S.splitBudget);
                nL = (uint32_t)lroundf(fmaxf(S.lightRatio * nB, 1.f));
            }
            const uint32_t nLv = first ? nL : (uint32_t)lroundf(fmaxf(S.lightRatio * 1u, 1.f));
            const F3 A = LightSamples(S, spec, k, len, sp, fr, wol, mat, nLv, first ? nB : 1u, lc, t.shadow);
            rr = (S.maxPL == 0) ? Clampf(RRProb(mat, wol), 0.0f, 0.997f) : 1.0f;
            if (first) {
                reflRoot = A;
                rSp = sp;
                rN = fr.z;
                rWol = wol;
                rMat = h.mat;
                rSides = mat.sides;
                rLcSum = lc.sum;
                rRR = rr;
                s = 0;
                wantRoot = true;
            } else {
                if (top >= 0) {
                    const uint32_t sl = Slot(top);
                    stA[sl][0] = topA.x;
                    stA[sl][1] = topA.y;
                    stA[sl][2] = topA.z;
                }
                ++top;
                topA = A;
                wantChain = true;
            }
        }
        // One call site of the BSDF sampler for both kinds of vertex.  A chain vertex whose sample
        // is a blocker ends its chain, and the first vertex's next sample follows in the same pass.
        const uint32_t nLc = (uint32_t)lroundf(fmaxf(S.lightRatio * 1u, 1.f));
        for (int pass = 0; pass < 2 && (wantRoot || wantChain); ++pass) {
            if (wantRoot && s >= nB) {
                L = em1 + reflRoot;
                return true;
            }
            F3 attCos;
            float den;
            bool cut;
            const DevMaterial vm = wantRoot ? LoadGlobal(S.mats + rMat) : mat;
            const Frame vfr = wantRoot ? FrameFromZ(rN) : fr;
            const bool ok = SampleAt(S, wantRoot ? key : k, wantRoot ? 1u : len, wantRoot ? 0u : branch, wantRoot ? s : 0u,
                                     wantRoot ? rSp : sp, vfr, wantRoot ? rWol : wol, vm, wantRoot ? nB : 1u,
                                     wantRoot ? nL : nLc, wantRoot ? rRR : rr, wantRoot ? rLcSum : lc.sum, attCos, den,
                                     cut);
            if (wantRoot) {
                if (ok) {
                    rAttCos = attCos;
                    rDen = den;
                    rCut = cut;
                    cutMask = 0;
                    return false;
                }
                ++s;                                        // a blocker: try the next sample
                --pass;
                continue;
            }
            if (ok) {
                const uint32_t sl = Slot(top);
                stW[sl][0] = attCos.x;
                stW[sl][1] = attCos.y;
                stW[sl][2] = attCos.z;
                stDen[sl] = den;
                cutMask = (cutMask & ~(1ull << sl)) | ((uint64_t)cut << sl);
                return false;
            }
            FoldChain(S);                                   // a blocker: this vertex has no subtree
            wantChain = false;
            wantRoot = true;
        }
        return false;
    }
    PG3_DEV F3 Result() const { return L; }
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
