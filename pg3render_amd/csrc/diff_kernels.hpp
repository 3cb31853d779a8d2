// Differentiable path tracing (-sb 1) with frozen sampling and path-replay backpropagation.
//
// The estimate of one sample is L = c0 + sum_v (N_v + E_v): N_v the MIS-weighted light sample of
// vertex v, E_v the MIS-weighted emission found by its BSDF sample, both carrying the throughput
// beta_v = prod_{u<v} w_u with w_u = f_u cos / pdf_u.  Sampling is frozen: directions, pdfs, MIS
// weights and light-picking probabilities are computed from the sampling parameters (mats, EM
// tables), and only the integrand values depend on the differentiated parameters (BSDF values f
// of Lambert / GGX-conductor materials read from evalMats, EM radiance read from evalTexels).
// The estimator is then a smooth function of the parameters whose expectation is the rendered
// image, and its derivative is an unbiased gradient of the image.
//
// Path replay (the adjoint recomputes each path from its (pixel, iteration) RNG stream instead of
// storing vertices): a primal replay gives L; a second replay walks the same vertices keeping the
// prefix P of L collected so far, and at vertex v adds
//   d N_v / d theta                                  (f_v at the light direction, Le of the EM)
//   (d w_v / d theta) / w_v * (L - P)                (every later term carries w_v once)
//   d E_v / d EM texels                              (emission found by the BSDF sample)
// each weighted by the adjoint of the pixel, dLoss/dI(pixel) / N.  Scalar parameters per material:
// Lambert albedo RGB (d log f_c / d rho_c = 1 / rho_c); GGX conductor tint RGB (1 / tint_c) and
// alpha (GgxCondDLogDAlpha).  EM texel gradients are added with FP64 atomics.
#pragma once

#include "pt_kernels.hpp"

namespace pg3 {

// EM-texel gradients are gathered per workgroup in an LDS cache of FP64 sums before they reach the
// global FP64 buffer: texel t owns slot t % kEmCacheSlots once a lane claims it (compare-and-swap
// on the slot key); a texel whose slot belongs to another texel goes straight to global memory.
// Importance sampling sends most light samples to a few bright texels, whose global atomics
// would otherwise serialise; the cache is flushed once per workgroup.
constexpr uint32_t kEmCacheSlots = 4096;

struct GradSink {
    F3 adj;            // dLoss/dI(pixel) / N, per channel
    F3 total;          // the sample's radiance from the primal replay
    double *em;        // EM texel gradients (w*h*3), FP64
    double *scal;      // the wavefront's LDS slice of scalar gradients (LDS)
    int32_t *cacheKeys;   // workgroup LDS cache (kEmCacheSlots keys, -1 = free), may be null
    double *cacheVals;    // kEmCacheSlots x 3 FP64 sums (LDS)
};

PG3_DEV void EmGradAdd(const GradSink &g, uint32_t t, F3 v) {
    if (g.cacheKeys) {
        const uint32_t s = t & (kEmCacheSlots - 1u);
        int32_t old = -1;           // the slot's key before the exchange
        __hip_atomic_compare_exchange_strong((__attribute__((address_space(3))) int32_t *)(g.cacheKeys + s), &old,
                                             (int32_t)t, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP);
        if (old == -1 || old == (int32_t)t) {
            LdsAdd(g.cacheVals + s * 3 + 0, (double)v.x);
            LdsAdd(g.cacheVals + s * 3 + 1, (double)v.y);
            LdsAdd(g.cacheVals + s * 3 + 2, (double)v.z);
            return;
        }
    }
    double *p = g.em + (size_t)t * 3;
    GlobalAdd(p + 0, (double)v.x);
    GlobalAdd(p + 1, (double)v.y);
    GlobalAdd(p + 2, (double)v.z);
}

PG3_DEV bool IsDiffMaterial(int32_t type) { return type == kMatLambert || type == kMatGgxCond; }

// Texel gradients of a term W * Le(footprint) (per channel).
PG3_DEV void ScatterEm(const GradSink &g, const EmFoot &f, F3 W) {
    const F3 a = g.adj * W;
    if (IsZero(a)) return;
    const float wa = (1 - f.xl) * (1 - f.yl), wb = f.xl * (1 - f.yl), wc = (1 - f.xl) * f.yl, wd = f.xl * f.yl;
    const uint32_t idx[4] = {f.a, f.b, f.c, f.d};
    const float wt[4] = {wa, wb, wc, wd};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (wt[k] == 0.f) continue;
        EmGradAdd(g, idx[k], a * wt[k]);
    }
}

// Scalar-parameter gradients of a term X whose value is proportional to f(wil, wol) of material m.
PG3_DEV void MatGrad(const GradSink &g, int32_t base, const DevMaterial &m, F3 wil, F3 wol, F3 X) {
    const F3 a = g.adj * X;
    if (m.type == kMatLambert) {
        if (m.diffuse.x > 0.f) LdsAdd(g.scal + base + 0, (double)(a.x / m.diffuse.x));
        if (m.diffuse.y > 0.f) LdsAdd(g.scal + base + 1, (double)(a.y / m.diffuse.y));
        if (m.diffuse.z > 0.f) LdsAdd(g.scal + base + 2, (double)(a.z / m.diffuse.z));
    } else if (m.type == kMatGgxCond) {
        if (m.glossy.x > 0.f) LdsAdd(g.scal + base + 0, (double)(a.x / m.glossy.x));
        if (m.glossy.y > 0.f) LdsAdd(g.scal + base + 1, (double)(a.y / m.glossy.y));
        if (m.glossy.z > 0.f) LdsAdd(g.scal + base + 2, (double)(a.z / m.glossy.z));
        const float s = a.x + a.y + a.z;
        if (s != 0.f) LdsAdd(g.scal + base + 3, (double)(s * GgxCondDLogDAlpha(m, wil, wol)));
    }
}

// One vertex of the differentiable estimator (the structure of ShadeVertex with frozen sampling).
// Adds this vertex's terms to ps.L (which is therefore the prefix P during the adjoint replay),
// and with kGrad their gradients.  Returns true if the path continues.
template <bool kGrad>
PG3_DEV bool DiffVertex(const SceneView &S, const LightSpec &spec, const RngKey &key, PathState &ps, bool hit,
                        const Hit &h, const GradSink &g, uint32_t &segments) {
    const uint32_t len = ps.len;
    const bool hasPrev = ps.flags & kHasPrev, prevFinite = ps.flags & kPrevFinite;
    if (!hit) {
        if (len >= S.minPL && S.bgLight >= 0) {                                       // :408-422
            F3 W;
            if (!hasPrev) {
                W = mk(1.f);
            } else if (prevFinite) {
                const float lp = BgPdf(S, spec, ps.dir, FrameFromZ(ps.prevN), ps.flags >> 4) *
                                 PrevPickProb(S, spec, key, ps, (uint32_t)S.bgLight);
                W = ps.eThr * MisPower(ps.prevBsdfPdf, lp);
            } else {
                W = ps.eThr;
            }
            if (spec.em != kEmAbsent) {
                const EmFoot fp = EmFootprint(S.em, ps.dir);
                ps.L += W * EmEvalFoot(S.evalTexels, fp);
                if (kGrad) ScatterEm(g, fp, W);
            } else {
                ps.L += W * LoadGlobal(S.lights + S.bgLight).radiance;
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

    VertexRng vr;
    RngVertex(key, len, vr);
    const uint32_t sides = mat.sides;
    const bool diffMat = IsDiffMaterial(mat.type);
    const bool separate = diffMat && (S.evalMats != S.mats);
    const DevMaterial matE = separate ? LoadGlobal(S.evalMats + h.mat) : mat;
    const int32_t gb = (kGrad && diffMat) ? ((GlobalPtr<int32_t>)AsGlobal(S.gradBase))[h.mat] : -1;
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
            const bool emLight = (li == S.bgLight) && (spec.em != kEmAbsent);
            EmFoot lf;
            float cf = 0.f;
            F3 Lc = ls.sample;
            if (emLight) {
                // radiance of the sampled direction from the evaluation texels, times the cosine
                // the sampler applies (zero on a side the material does not sample)
                const float c = Dot(ls.wig, fr.z);
                cf = (((sides & kSideFront) && c > 0.f) || ((sides & kSideBack) && c < 0.f)) ? fabsf(c) : 0.f;
                lf = EmFootprint(S.em, ls.wig);
                Lc = EmEvalFoot(S.evalTexels, lf) * cf;
            }
            if (MaxC(Lc) > 0.f || (kGrad && cf > 0.f)) {
                const F3 wilL = ToLocal(fr, ls.wig);
                BsdfRec r;
                EvalBsdf(mat, wilL, wol, r);
                F3 f = r.att;
                if (separate) {
                    BsdfRec re;
                    EvalBsdf(matE, wilL, wol, re);
                    f = re.att;
                }
                F3 K;
                if (ls.pdfW != kInf) {
                    const float lightPdf = ls.pdfW * ls.prob;
                    K = ps.beta * (MisPower(lightPdf, r.pdfW * r.compProb) / lightPdf);
                } else {
                    K = ps.beta / ls.prob;
                }
                if (!Occluded(S, sp, ls.wig, ls.dist)) {
                    const F3 contrib = K * Lc * f;
                    ps.L += contrib;
                    if (kGrad) {
                        if (emLight) ScatterEm(g, lf, K * f * cf);
                        if (gb >= 0) MatGrad(g, gb, matE, wilL, wol, contrib);
                    }
                }
            }
        }
    }
    BsdfRec r;                                                                        // :267-303
    SampleBsdf(mat, wol, vr.u[kDimBsdf0], vr.u[kDimBsdf1], vr.u[kDimBsdf2], r);
    if (IsBlocker(r)) return false;
    const float cosAbs = fabsf(r.wil.z);
    const bool finite = (r.pdfW != kInf);
    F3 att = r.att;
    if (diffMat && finite) {
        BsdfRec re;
        EvalBsdf(matE, r.wil, wol, re);
        att = re.att;
    }
    if (kGrad && gb >= 0 && finite) MatGrad(g, gb, matE, r.wil, wol, g.total - ps.L);
    if (finite) {
        ps.prevBsdfPdf = r.pdfW * r.compProb;
        ps.eThr = ps.beta * att * cosAbs / ps.prevBsdfPdf;
        ps.beta = ps.beta * ((att * cosAbs) / (r.pdfW * r.compProb));
    } else {
        ps.eThr = ps.beta * att / r.compProb;
        ps.beta = ps.beta * (att / r.compProb);
    }
    ps.flags = kHasPrev | (finite ? kPrevFinite : 0u) | kComputeRefl | (sides << 4);
    ps.prevN = fr.z;
    ps.lcSum = lc.sum;
    ps.org = sp;
    ps.dir = ToWorld(fr, r.wil);
    ps.tmin = EpsRayCos(cosAbs);
    ps.len = len + 1;
    return true;
}

// A work item (8x8 tile x iterations) of the differentiable forward (kAdjoint false: per-pixel
// sums into fb) or of the adjoint (kAdjoint true: each sample is traced twice, the primal replay
// giving its radiance and the second replay the gradients).  Samples are handed out in order, one
// to each lane whose path has ended.  sampleL (may be null) holds one radiance per sample of the
// launch, index (iteration - firstIter) * sampleStride + pixel: the forward writes it, and the
// adjoint then reads the radiance instead of running the primal replay.  view (may be null): the
// camera of the tile's view in a multi-view scene; fb, adjImg and sampleL then point at the view's
// part of buffers that hold every view (sampleStride = views x pixels).
template <bool kAdjoint>
PG3_DEV void TraceTileDiff(const SceneView &S, const LightSpec &spec, uint32_t x0, uint32_t y0, uint32_t itBegin,
                           uint32_t itEnd, double *acc, double *fb, const float *adjImg, float invN, double *gradEm,
                           double *gradScal, float *sampleL, uint32_t firstIter, int32_t *cacheKeys,
                           double *cacheVals, uint32_t &segments, uint32_t &samples, const DevView *view,
                           size_t sampleStride) {
    const uint32_t lane = __lane_id();
    if (!kAdjoint) {
        acc[lane * 3 + 0] = 0.0;
        acc[lane * 3 + 1] = 0.0;
        acc[lane * 3 + 2] = 0.0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    const uint32_t total = 64u * (itEnd - itBegin);
    uint32_t next = 0;
    bool busy = false, second = false;
    uint32_t pix = 0, pixel = 0, iter = 0;
    RngKey key{};
    PathState ps;
    GradSink g;
    g.adj = mk(0.f);
    g.total = mk(0.f);
    g.em = gradEm;
    g.scal = gradScal;
    g.cacheKeys = cacheKeys;
    g.cacheVals = cacheVals;
    for (;;) {
        const uint64_t idle = __ballot(!busy);
        if (next < total && idle != 0) {
            const uint32_t q = next + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
            if (!busy && q < total) {
                pix = q & 63u;
                const uint32_t x = x0 + (pix & 7u), y = y0 + (pix >> 3);
                if (x < S.resX && y < S.resY) {
                    pixel = y * S.resX + x;
                    iter = itBegin + (q >> 6);
                    F3 org, dir;
                    CameraSample(S, view, pixel, iter, key, org, dir);
                    InitPath(ps, org, dir);
                    busy = true;
                    second = false;
                    if (kAdjoint) {
                        const GlobalPtr<float> a = AsGlobal(adjImg + (size_t)pixel * 3);
                        g.adj = mk(a[0], a[1], a[2]) * invN;
                        if (sampleL) {
                            const size_t si = ((size_t)(iter - firstIter) * sampleStride + pixel) * 3;
                            const GlobalPtr<float> l = AsGlobal(sampleL + si);
                            g.total = mk(l[0], l[1], l[2]);
                            second = true;
                        }
                    }
                }
            }
            next += (uint32_t)__popcll(idle);
        }
        if (__ballot(busy) == 0) {
            if (next >= total) break;
            continue;
        }
        if (busy) {
            Hit h;
            h.dist = 1e36f;
            const bool hit = Intersect(S, ps.org, ps.dir, ps.tmin, h);
            bool cont;
            if (kAdjoint && second) cont = DiffVertex<true>(S, spec, key, ps, hit, h, g, segments);
            else cont = DiffVertex<false>(S, spec, key, ps, hit, h, g, segments);
            if (!cont) {
                if (!kAdjoint) {
                    if (sampleL) {
                        float *l = sampleL + ((size_t)(iter - firstIter) * sampleStride + pixel) * 3;
                        l[0] = ps.L.x;
                        l[1] = ps.L.y;
                        l[2] = ps.L.z;
                    }
                    atomicAdd(acc + pix * 3 + 0, (double)ps.L.x);
                    atomicAdd(acc + pix * 3 + 1, (double)ps.L.y);
                    atomicAdd(acc + pix * 3 + 2, (double)ps.L.z);
                    ++samples;
                    busy = false;
                } else if (!second) {
                    g.total = ps.L;              // primal replay done: replay the same path again
                    F3 org, dir;
                    CameraSample(S, view, pixel, iter, key, org, dir);
                    InitPath(ps, org, dir);
                    second = true;
                } else {
                    ++samples;
                    busy = false;
                }
            }
        }
    }
    if (!kAdjoint) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const uint32_t x = x0 + (lane & 7u), y = y0 + (lane >> 3);
        if (fb && x < S.resX && y < S.resY) {
            double *p = fb + (size_t)(y * S.resX + x) * 3;
            atomicAdd(p + 0, acc[lane * 3 + 0]);
            atomicAdd(p + 1, acc[lane * 3 + 1]);
            atomicAdd(p + 2, acc[lane * 3 + 2]);
        }
    }
}

}  // namespace pg3
