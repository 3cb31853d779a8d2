// Device math of the gfx950 path-tracing hot path: vectors, frames, the Philox4x32-10 counter RNG,
// sampling warps, Fresnel terms, GGX microfacet functions and the BSDFs.  Every function follows
// the cited PG3Render file:line in behaviour; all arithmetic is FP32 as in the reference.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PG3_DEV __device__ __forceinline__
#define PG3_HD __host__ __device__ __forceinline__

namespace pg3 {

constexpr float kPiF = 3.14159265358979f;          // math.hxx:12
constexpr float kPiInvF = 1.f / 3.14159265358979f;  // math.hxx:13
constexpr float kTinyF = 9.5367431640625E-7f;       // math.hxx:45
constexpr float kEpsRay = 1e-3f;                    // geom.hxx:16
constexpr float kEpsDist = 1e-6f;                   // geom.hxx:27
constexpr float kBlockerEps = 1e-5f;                // materials.hxx:16
constexpr float kInf = __builtin_huge_valf();

PG3_HD bool IsTiny(float x) { return fabsf(x) < kTinyF; }
PG3_HD float Sqr(float x) { return x * x; }
PG3_HD float SafeSqrt(float x) { return sqrtf(fmaxf(0.f, x)); }
PG3_HD float Clampf(float x, float a, float b) { return fminf(fmaxf(x, a), b); }

struct F3 {
    float x, y, z;
};
PG3_HD F3 mk(float x, float y, float z) { return F3{x, y, z}; }
PG3_HD F3 mk(float a) { return F3{a, a, a}; }
PG3_HD F3 operator+(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
PG3_HD F3 operator-(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
PG3_HD F3 operator*(F3 a, F3 b) { return F3{a.x * b.x, a.y * b.y, a.z * b.z}; }
PG3_HD F3 operator*(F3 a, float s) { return F3{a.x * s, a.y * s, a.z * s}; }
PG3_HD F3 operator*(float s, F3 a) { return F3{a.x * s, a.y * s, a.z * s}; }
PG3_HD F3 operator/(F3 a, float s) { return F3{a.x / s, a.y / s, a.z / s}; }
PG3_HD F3 operator-(F3 a) { return F3{-a.x, -a.y, -a.z}; }
PG3_HD F3 &operator+=(F3 &a, F3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
PG3_HD float Dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PG3_HD F3 Cross(F3 a, F3 b) { return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
PG3_HD float LenSqr(F3 a) { return Dot(a, a); }
PG3_HD float Length(F3 a) { return sqrtf(Dot(a, a)); }
PG3_HD F3 Normalize(F3 a) { return a / sqrtf(Dot(a, a)); }            // math.hxx:653-658
// Division by the hardware reciprocal: a * v_rcp_f32(b), within 2 ulp of the correctly rounded
// quotient (1 ulp reciprocal, then a rounded product); one transcendental and one multiply in
// place of the ~10-instruction correctly rounded sequence.  (`__fdividef` would not do: this HIP
// defines it as plain `/`.)  Used where a one-ulp change only moves a value by ~1e-7 relative:
// pdf ratios, direction normalisation, MIS weights, path throughput and contributions.  BSDF
// values and pdfs, and the sampling warps whose output a one-ulp change would move by more (the
// steerable tree descent rescales its random number at every level), keep correctly rounded
// division.
#if defined(__HIP_DEVICE_COMPILE__)
PG3_HD float FastDiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
#else
PG3_HD float FastDiv(float a, float b) { return a / b; }
#endif
PG3_HD F3 FastDiv(F3 a, float b) {
    const float r = FastDiv(1.0f, b);
    return F3{a.x * r, a.y * r, a.z * r};
}
PG3_HD float MaxC(F3 a) { return fmaxf(fmaxf(a.x, a.y), a.z); }        // math.hxx:535
PG3_HD bool IsZero(F3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
PG3_HD float Luminance(F3 a) { return 0.212671f * a.x + 0.715160f * a.y + 0.072169f * a.z; }  // spectrum.hxx:74-80

// Frame::SetFromZ / ToWorld / ToLocal, math.hxx:980-1000
struct Frame {
    F3 x, y, z;
};
PG3_HD Frame FrameFromZ(F3 z) {
    Frame f;
    f.z = z;
    const F3 tmp = (fabsf(z.x) > 0.99f) ? mk(0.f, 1.f, 0.f) : mk(1.f, 0.f, 0.f);
    f.y = Normalize(Cross(z, tmp));
    f.x = Cross(f.y, z);
    return f;
}
PG3_HD F3 ToWorld(const Frame &f, F3 a) { return f.x * a.x + f.y * a.y + f.z * a.z; }
PG3_HD F3 ToLocal(const Frame &f, F3 a) { return mk(Dot(a, f.x), Dot(a, f.y), Dot(a, f.z)); }

// ---------------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy as 1, 2, 3", SC'11).
// The reference draws sequentially from std::mt19937_64 (rng.hxx:16-61); the GPU path and the
// CPU oracle instead give every random number a fixed counter so paths can be traced in any
// order and replayed by the adjoint.
struct U4 {
    uint32_t a, b, c, d;
};
PG3_HD U4 Philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return U4{c0, c1, c2, c3};
}
PG3_HD float ToUnit(uint32_t v) { return (float)(v >> 8) * (1.0f / 16777216.0f); }

// Dimension slots of one path vertex (shared with the CPU oracle).
enum : uint32_t {
    kDimLightPick = 0, kDimLight0 = 1, kDimLight1 = 2, kDimLight2 = 3,
    kDimRR = 4, kDimBsdf0 = 5, kDimBsdf1 = 6, kDimBsdf2 = 7,
    kDimEstimateBase = 8,
};

// Random numbers of one vertex: block 0 holds dims 0-3, block 1 dims 4-7.
struct VertexRng {
    float u[8];
};
struct RngKey {
    uint32_t k0, k1, pixel, iteration, c3;   // c3 = scene << 16 | branch
};
PG3_HD float RngGet(const RngKey &k, uint32_t vertex, uint32_t sub, uint32_t dim) {
    const uint32_t c2 = (vertex << 16) | ((sub & 0xFFu) << 8) | (dim >> 2);
    const U4 r = Philox(k.pixel, k.iteration, c2, k.c3, k.k0, k.k1);
    const uint32_t w = (dim & 3) == 0 ? r.a : (dim & 3) == 1 ? r.b : (dim & 3) == 2 ? r.c : r.d;
    return ToUnit(w);
}
PG3_HD void RngVertex(const RngKey &k, uint32_t vertex, VertexRng &v) {
    const uint32_t c2 = (vertex << 16);
    const U4 a = Philox(k.pixel, k.iteration, c2, k.c3, k.k0, k.k1);
    const U4 b = Philox(k.pixel, k.iteration, c2 | 1u, k.c3, k.k0, k.k1);
    v.u[0] = ToUnit(a.a); v.u[1] = ToUnit(a.b); v.u[2] = ToUnit(a.c); v.u[3] = ToUnit(a.d);
    v.u[4] = ToUnit(b.a); v.u[5] = ToUnit(b.b); v.u[6] = ToUnit(b.c); v.u[7] = ToUnit(b.d);
}

// ---------------------------------------------------------------------------------------------
// geom.hxx helpers
PG3_HD float EpsRayCos(float c) { return (1.f + 2.f * (1.f - c)) * kEpsRay; }   // :23-26
PG3_HD F3 ReflectLocal(F3 d) { return mk(-d.x, -d.y, d.z); }                     // :108-111
PG3_HD float TanThetaSqr(F3 d) {                                                   // :96-106
    const float c2 = d.z * d.z;
    const float s2 = 1.f - c2;
    return (s2 <= 0.0f) ? 0.0f : s2 / c2;
}
PG3_HD F3 Reflect(F3 in, F3 n, bool &above) {                                       // :119-133
    const float d = Dot(in, n);
    above = d > 0.0f;
    return (2.0f * d) * n - in;
}
PG3_HD F3 Refract(F3 in, F3 n, float eta, bool &inAbove) {                          // :146-178
    const float ci = Dot(in, n);
    inAbove = (ci > 0.0f);
    if (inAbove) eta = 1.0f / eta;
    const float ct2 = 1 - (1 - ci * ci) * (eta * eta);
    if (ct2 < 0.0f) { bool d; return Reflect(in, n, d); }
    float ct = sqrtf(ct2);
    ct = (ci > 0.0f) ? -ct : ct;
    return n * (ci * eta + ct) - in * eta;
}
PG3_HD F3 LatLong2Dir(float u, float v) {                                           // :65-76
    const float phi = -(u - 0.5f) * 2 * kPiF;
    const float theta = v * kPiF;
    float st, ct, sp, cp;
    sincosf(theta, &st, &ct);
    sincosf(phi, &sp, &cp);
    return mk(st * cp, st * sp, ct);
}
PG3_HD void Dir2LatLong(F3 d, float &u, float &v) {                                 // :80-95
    const float phi = -atan2f(d.y, d.x);
    const float theta = acosf(d.z);
    u = Clampf(0.5f + phi * 0.5f * kPiInvF, 0.f, 1.f);
    v = Clampf(theta * kPiInvF, 0.f, 1.0f);
}

// sampling.hxx warps
// The pdf of a sample w is w.z / pi (the reference's oPdfW output); callers that need it compute
// it from the returned direction (an optional pdf pointer, null at some call sites, kept the
// BSDF record of the shading code in scratch memory after the compiler merged the calls).
PG3_HD F3 SampleCosHemisphereW(float u0, float u1) {                               // :289-307
    const float t1 = 2.f * kPiF * u0;
    const float t2 = sqrtf(1.f - u1);
    float s, c;
    sincosf(t1, &s, &c);
    return mk(c * t2, s * t2, sqrtf(u1));
}
PG3_HD float CosHemispherePdfW(float z) { return fmaxf(0.f, z) * kPiInvF; }        // :309-312
PG3_HD F3 SampleCosSphereParamPdfW(float u0, float u1, float u2, bool up, bool low, float &pdf) {  // :326-345
    F3 w = SampleCosHemisphereW(u0, u1);
    pdf = w.z * kPiInvF;
    if (up && low) { if (u2 < 0.5f) w = w * -1.0f; pdf *= 0.5f; }
    else if (low) w = w * -1.0f;
    return w;
}
PG3_HD float CosSpherePdfW(bool up, bool low, F3 d) {                               // :347-364
    if ((!up && d.z > 0.f) || (!low && d.z < 0.f)) return 0.f;
    float p = CosHemispherePdfW(fabsf(d.z));
    if (up && low) p *= 0.5f;
    return p;
}
PG3_HD F3 SamplePowerCosHemisphereW(float u0, float u1, float power) {              // :250-265
    const float t1 = 2.f * kPiF * u0;
    const float t2 = powf(u1, 1.f / (power + 1.f));
    const float t3 = sqrtf(1.f - t2 * t2);
    float s, c;
    sincosf(t1, &s, &c);
    return mk(c * t3, s * t3, t2);
}
PG3_HD float PowerCosHemispherePdfW(F3 d, float power) {                             // :276-284
    const float c = fmaxf(0.f, d.z);
    return (power + 1.f) * powf(c, power) * (kPiInvF * 0.5f);
}
PG3_HD F3 SampleUniformTriangle(F3 p0, F3 p1, F3 p2, float u0, float u1) {          // sampling.hxx:66-75
    const float xs = sqrtf(u0);
    const float b0 = 1.f - xs, b1 = u1 * xs;
    const float b2 = 1.0f - b0 - b1;
    return p0 * b0 + p1 * b1 + p2 * b2;
}

// physics.hxx
PG3_HD float FresnelDielectric(float ci, float eta) {                               // :5-47
    if (eta < 0) return 1.f;
    float we;
    if (ci < 0.f) { ci = -ci; we = eta; } else we = 1.f / eta;
    const float st2 = Sqr(we) * (1.f - Sqr(ci));
    const float ct = SafeSqrt(1.f - st2);
    if (st2 > 1.0f) return 1.0f;
    const float t2 = we * ci;
    const float rs = Sqr((t2 - ct) / (t2 + ct));
    const float t1 = we * ct;
    const float rp = Sqr((ci - t1) / (ci + t1));
    return 0.5f * (rp + rs);
}
PG3_HD float FresnelConductor(float ci, float eta, float k) {                       // :50-141
    if (ci < -0.00001f) return 0.0f;
    ci = Clampf(ci, 0.0f, 1.0f);
    const float c2 = ci * ci, s2 = 1 - c2, s4 = s2 * s2;
    const float e2 = eta * eta, k2 = k * k;
    const float t = e2 - k2 - s2;
    const float a2pb2 = SafeSqrt(t * t + 4 * k2 * e2);
    const float a = SafeSqrt(0.5f * (a2pb2 + t));
    const float term1 = a2pb2 + c2, term2 = 2 * a * ci;
    const float rs = (term1 - term2) / (term1 + term2);
    const float term3 = a2pb2 * c2 + s4, term4 = term2 * s2;
    const float rp = rs * (term3 - term4) / (term3 + term4);
    return 0.5f * (rp + rs);
}

// microfacet.hxx
PG3_HD float ReflectionJacobian(F3 wo, F3 h) {                                       // :16-31
    return 1.0f / (4.0f * fmaxf(Dot(h, wo), 0.000001f));
}
PG3_HD float RefractionJacobian(F3 wi, F3 wo, F3 h, float etaOutIn) {                // :34-56
    const float cOM = Dot(h, wo), cIM = Dot(h, wi);
    return fabsf(cOM) / fmaxf(Sqr(etaOutIn * cIM + cOM), 0.000001f);
}
PG3_HD F3 HalfwayReflection(F3 wi, F3 wo) {                                          // :60-84
    F3 h = wi + wo;
    const float len = Length(h);
    h = IsTiny(len) ? mk(0.f, 0.f, 1.f) : h / len;
    return h * ((h.z >= 0.0f) ? 1.0f : -1.0f);
}
PG3_HD F3 HalfwayRefraction(F3 wi, F3 wo, float etaInOut) {                         // :88-123
    F3 h = wi + wo * etaInOut;
    const float len = Length(h);
    h = IsTiny(len) ? wi : h / len;
    return h * ((h.z >= 0.0f) ? 1.0f : -1.0f);
}
PG3_HD float DistributionGgx(F3 m, float alpha) {                                    // :505-527
    if (m.z <= 0.f) return 0.0f;
    const float a2 = alpha * alpha;
    const float c2 = m.z * m.z;
    const float t = c2 * (a2 + TanThetaSqr(m));
    return a2 / (kPiF * t * t);
}
PG3_HD float SmithG1(F3 v, F3 m, float alpha) {                                      // :529-554
    if (m.z <= 0) return 0.0f;
    if ((v.z * Dot(v, m)) < 0.0f) return 0.0f;
    const float root = sqrtf(1.0f + alpha * alpha * TanThetaSqr(v));
    return 2.0f / (1.0f + root);
}
PG3_HD void SampleGgxP11(float thetaI, float ux, float uy, float &sx, float &sy) {   // :605-669
    if (thetaI < 0.0001f) {
        const float x = fminf(ux, 0.9999f);
        const float r = SafeSqrt(x / (1 - x));
        float s, c;
        sincosf(2 * kPiF * uy, &s, &c);
        sx = r * c;
        sy = r * s;
        return;
    }
    const float tanI = tanf(thetaI);
    const float tanIInv = 1.0f / tanI;
    const float G1 = 2.0f / (1.0f + SafeSqrt(1.0f + 1.0f / (tanIInv * tanIInv)));
    float A = 2.0f * ux / G1 - 1.0f;
    if (fabsf(A) == 1.0f) A -= (signbit(A) ? -1.0f : 1.0f) * 1e-4f;
    const float B = tanI;
    const float tmp = 1.0f / (A * A - 1.0f);
    const float D = SafeSqrt(B * B * tmp * tmp - (A * A - B * B) * tmp);
    const float x1 = B * tmp - D, x2 = B * tmp + D;
    sx = (A < 0.0f || x2 > (1.0f / tanI)) ? x1 : x2;
    float ys, yh;
    if (uy > 0.5f) { ys = 1.0f; yh = 2.0f * (uy - 0.5f); }
    else { ys = -1.0f; yh = 2.0f * (0.5f - uy); }
    // rational fit of the inverse conditional CDF used by the reference (microfacet.hxx:655-662)
    const float z =
        (yh * (yh * (yh * -0.365728915865723f + 0.790235037209296f) - 0.424965825137544f) +
         0.000152998850436920f) /
        (yh * (yh * (yh * (yh * 0.169507819808272f - 0.397203533833404f) - 0.232500544458471f) + 1.0f) -
         0.539825872510702f);
    sy = ys * z * sqrtf(1.0f + sx * sx);
}
PG3_HD F3 SampleGgxVisibleNormals(F3 wo, float alpha, float ux, float uy) {         // :678-729
    const F3 ws = Normalize(mk(wo.x * alpha, wo.y * alpha, fmaxf(wo.z, 0.0f)));
    float theta = 0.0f, phi = 0.0f;
    if (ws.z < 0.999f) { theta = acosf(ws.z); phi = atan2f(ws.y, ws.x); }
    float sp, cp;
    sincosf(phi, &sp, &cp);
    float sx, sy;
    SampleGgxP11(theta, ux, uy, sx, sy);
    const float rx = (sx * cp - sy * sp) * alpha;
    const float ry = (sx * sp + sy * cp) * alpha;
    const float inv = 1.0f / sqrtf(rx * rx + ry * ry + 1.0f);
    return mk(-rx * inv, -ry * inv, inv);
}
PG3_HD float GgxVndfPdf(F3 wo, F3 h, float D, float alpha) {                        // :732-756
    if (h.z <= 0.f) return 0.0f;
    return (SmithG1(wo, h, alpha) * fabsf(Dot(h, wo)) * D) / fmaxf(wo.z, 0.00001f);
}

// ---------------------------------------------------------------------------------------------
// Materials (materials.hxx), stored pre-digested as the constructors leave them.
enum : int32_t { kMatLambert = 0, kMatPhong = 1, kMatGgxCond = 2, kMatGgxDiel = 3, kMatSmoothCond = 4, kMatSmoothDiel = 5 };
enum : uint32_t { kSideFront = 1u, kSideBack = 2u };

struct DevMaterial {
    int32_t type;
    int32_t lightId;
    uint32_t sides;        // MaterialProperties light-sampling sides
    float phongExp;
    F3 diffuse;            // Lambert/Phong diffuse (Phong already halved when glossy)
    F3 glossy;             // Phong glossy reflectance or GGX tint
    float alpha, eta, etaInv, k;
};

struct BsdfRec {            // MaterialRecord, materials.hxx:48-207
    F3 wil, att;
    float pdfW, compProb;
};
PG3_HD bool IsBlocker(const BsdfRec &r) { return MaxC(r.att) * fabsf(r.wil.z) <= kBlockerEps; }

PG3_HD float PhongPdf(const DevMaterial &m, F3 wol, F3 wil, float dEst, float gEst) {  // :506-538
    const float tot = dEst + gEst;
    if (tot < kBlockerEps) return CosHemispherePdfW(wil.z);
    const Frame lf = FrameFromZ(ReflectLocal(wol));
    const F3 wc = ToLocal(lf, wil);
    return (dEst / tot) * CosHemispherePdfW(wil.z) + (gEst / tot) * PowerCosHemispherePdfW(wc, m.phongExp);
}
PG3_HD F3 PhongEval(const DevMaterial &m, F3 wil, F3 wol) {                          // :381-422
    if (wil.z <= 0.f || wol.z <= 0.f) return mk(0.f);
    const float cc = (m.phongExp + 2.0f) / (2.0f * kPiF);
    const float c = fmaxf(Dot(ReflectLocal(wil), wol), 0.f);
    return m.diffuse / kPiF + m.glossy * (cc * powf(c, m.phongExp));
}
PG3_HD F3 GgxCondEval(const DevMaterial &m, F3 wil, F3 wol, F3 h, float D, float F) {  // :909-938
    if (wil.z <= 0.f || wol.z <= 0.f || IsTiny(4.0f * wil.z * wol.z)) return mk(0.f);
    const float G = SmithG1(wil, h, m.alpha) * SmithG1(wol, h, m.alpha);
    return m.glossy * ((F * G * D) / (4.0f * wil.z * wol.z));
}
// d log f / d alpha of the GGX conductor value at (wil, wol) (f = tint F G1(i) G1(o) D / (4 ci co),
// alpha in D and both G1): d log D = 2/a - 4a / (a^2 + tan^2(h)), d log G1(v) = -a tan^2(v) /
// (R (1 + R)) with R = sqrt(1 + a^2 tan^2(v)).  Zero where the value is zero.
PG3_HD float GgxCondDLogDAlpha(const DevMaterial &m, F3 wil, F3 wol) {
    if (wil.z <= 0.f || wol.z <= 0.f) return 0.f;
    const F3 h = HalfwayReflection(wil, wol);
    if (h.z <= 0.f || (wil.z * Dot(wil, h)) < 0.f || (wol.z * Dot(wol, h)) < 0.f) return 0.f;
    const float a = m.alpha, a2 = a * a;
    float d = 2.f / a - 4.f * a / (a2 + TanThetaSqr(h));
    const float ti = TanThetaSqr(wil), to = TanThetaSqr(wol);
    const float ri = sqrtf(1.f + a2 * ti), ro = sqrtf(1.f + a2 * to);
    d -= a * ti / (ri * (1.f + ri));
    d -= a * to / (ro * (1.f + ro));
    return d;
}
struct DielCtx {
    F3 wolS, wilS, h;
    bool below, refl;
    float etaS, etaInvS, F, D;
};
PG3_HD F3 DielEval(const DevMaterial &m, const DielCtx &c) {                         // :1169-1225
    const float ci = fabsf(c.wilS.z), co = fabsf(c.wolS.z);
    if ((!c.refl && IsTiny(ci * co)) || (c.refl && IsTiny(4.0f * ci * co))) return mk(0.f);
    const float G = SmithG1(c.wilS, c.h, m.alpha) * SmithG1(c.wolS, c.h, m.alpha);
    float v;
    if (c.refl) {
        v = (c.F * G * c.D) / (4.0f * ci * co);
    } else {
        const float cMI = Dot(c.h, c.wilS), cMO = Dot(c.h, c.wolS);
        v = ((fabsf(cMI) * fabsf(cMO)) / (ci * co)) *
            ((Sqr(c.etaInvS) * (1.0f - c.F) * G * c.D) / Sqr(cMI + c.etaInvS * cMO));
        v *= Sqr(c.etaS);
    }
    return mk(v);
}
PG3_HD float DielPdf(const DevMaterial &m, const DielCtx &c) {                       // :1227-1266
    if (c.wolS.z < 0.f) return 0.0f;
    const float vn = GgxVndfPdf(c.wolS, c.h, c.D, m.alpha);
    const float jac = c.refl ? ReflectionJacobian(c.wilS, c.h) : RefractionJacobian(c.wolS, c.wilS, c.h, c.etaInvS);
    return vn * jac * (c.refl ? c.F : (1.f - c.F));
}

// AbstractMaterial::EvalBsdf with kOptSamplingProbs (pdfW, compProb)
PG3_HD void EvalBsdf(const DevMaterial &m, F3 wil, F3 wol, BsdfRec &r) {
    // the record is written once at each exit: its fields stay in registers (a record written
    // field by field in every case of the switch was kept in scratch memory)
    F3 o_wil = mk(0.f), o_att = mk(0.f);
    float o_pdf = 0.f, o_comp = 0.f;
    o_wil = wil;
    switch (m.type) {
    case kMatLambert:                                                                 // :268-345
        o_att = (wil.z <= 0.f || wol.z <= 0.f) ? mk(0.f) : m.diffuse / kPiF;
        o_pdf = CosHemispherePdfW(wil.z);
        o_comp = 1.f;
        break;
    case kMatPhong: {                                                                 // :424-437, 579-596
        o_att = PhongEval(m, wil, wol);
        const float dEst = Luminance(m.diffuse);
        const float gEst = Luminance(m.glossy) * (0.5f + 0.5f * fmaxf(wol.z, 0.f));
        o_pdf = PhongPdf(m, wol, wil, dEst, gEst);
        o_comp = 1.0f;
        break;
    }
    case kMatGgxCond: {                                                               // :808-816, 889-959
        const F3 h = HalfwayReflection(wil, wol);
        const float D = DistributionGgx(h, m.alpha);
        const float F = FresnelConductor(Dot(h, wol), m.eta, m.k);
        o_att = GgxCondEval(m, wil, wol, h, D, F);
        o_comp = 1.0f;
        o_pdf = (wol.z < 0.f) ? 0.0f : GgxVndfPdf(wol, h, D, m.alpha) * ReflectionJacobian(wol, h);
        break;
    }
    case kMatGgxDiel: {                                                               // :1023-1031, 1134-1167
        DielCtx c;
        c.below = (wol.z < 0.f);
        c.refl = ((wil.z > 0.f) == (wol.z > 0.f));
        const float s = c.below ? -1.0f : 1.0f;
        c.wolS = wol * s;
        c.wilS = wil * s;
        c.etaS = c.below ? m.etaInv : m.eta;
        c.etaInvS = c.below ? m.eta : m.etaInv;
        c.h = c.refl ? HalfwayReflection(c.wilS, c.wolS) : HalfwayRefraction(c.wilS, c.wolS, c.etaInvS);
        c.D = DistributionGgx(c.h, m.alpha);
        c.F = FresnelDielectric(Dot(c.h, c.wolS), c.etaS);
        o_att = DielEval(m, c);
        o_pdf = DielPdf(m, c);
        o_comp = 1.0f;
        break;
    }
    default:                                                                          // :612-654
        o_att = mk(0.f);
        o_pdf = 0.f;
        o_comp = 0.f;
        break;
    }
    r = BsdfRec{o_wil, o_att, o_pdf, o_comp};
}

// AbstractMaterial::SampleBsdf with explicit uniforms (u0,u1 direction, u2 component choice)
PG3_HD void SampleBsdf(const DevMaterial &m, F3 wol, float u0, float u1, float u2, BsdfRec &r) {
    // the record is written once at each exit: its fields stay in registers (a record written
    // field by field in every case of the switch was kept in scratch memory)
    F3 o_wil = mk(0.f), o_att = mk(0.f);
    float o_pdf = 0.f, o_comp = 0.f;
    switch (m.type) {
    case kMatLambert:                                                                 // :285-292
        o_wil = SampleCosHemisphereW(u0, u1);
        o_pdf = o_wil.z * kPiInvF;
        o_att = (o_wil.z <= 0.f || wol.z <= 0.f) ? mk(0.f) : m.diffuse / kPiF;
        o_comp = 1.f;
        break;
    case kMatPhong: {                                                                 // :442-504
        const float dEst = Luminance(m.diffuse);
        const float gEst = Luminance(m.glossy) * (0.5f + 0.5f * fmaxf(wol.z, 0.f));
        const float tot = dEst + gEst;
        if (tot < kBlockerEps) {
            o_att = mk(0.f);
            o_wil = SampleCosHemisphereW(u0, u1);
            o_pdf = o_wil.z * kPiInvF;
            o_comp = 1.f;
            r = BsdfRec{o_wil, o_att, o_pdf, o_comp};
            return;
        }
        if (u2 * tot < dEst) {
            o_wil = SampleCosHemisphereW(u0, u1);
        } else {
            const F3 cs = SamplePowerCosHemisphereW(u0, u1, m.phongExp);
            o_wil = ToWorld(FrameFromZ(ReflectLocal(wol)), cs);
        }
        o_comp = 1.f;
        o_pdf = PhongPdf(m, wol, o_wil, dEst, gEst);
        o_att = (o_wil.z > 0.0f) ? PhongEval(m, o_wil, wol) : mk(0.f);
        break;
    }
    case kMatGgxCond: {                                                               // :819-849
        const F3 h = SampleGgxVisibleNormals(wol, m.alpha, u0, u1);
        const float D = DistributionGgx(h, m.alpha);
        bool above;
        o_wil = Reflect(wol, h, above);
        const float F = FresnelConductor(Dot(h, wol), m.eta, m.k);
        o_comp = 1.0f;
        o_pdf = (wol.z < 0.f) ? 0.0f : GgxVndfPdf(wol, h, D, m.alpha) * ReflectionJacobian(wol, h);
        o_att = (!above || o_wil.z < 0.0f) ? mk(0.f) : GgxCondEval(m, o_wil, wol, h, D, F);
        break;
    }
    case kMatGgxDiel: {                                                               // :1034-1089
        DielCtx c;
        c.below = (wol.z < 0.f);
        const float s = c.below ? -1.0f : 1.0f;
        c.wolS = wol * s;
        c.etaS = c.below ? m.etaInv : m.eta;
        c.etaInvS = c.below ? m.eta : m.etaInv;
        c.h = SampleGgxVisibleNormals(c.wolS, m.alpha, u0, u1);
        c.D = DistributionGgx(c.h, m.alpha);
        c.F = FresnelDielectric(Dot(c.h, c.wolS), c.etaS);
        bool above;
        float cosAbs;
        if (u2 <= c.F) {
            c.wilS = Reflect(c.wolS, c.h, above);
            cosAbs = c.wilS.z;
            c.refl = true;
        } else {
            c.wilS = Refract(c.wolS, c.h, c.etaS, above);
            cosAbs = -c.wilS.z;
            c.refl = false;
        }
        o_wil = c.wilS * s;
        o_pdf = DielPdf(m, c);
        o_comp = 1.0f;
        o_att = (!above || cosAbs < 0.0f) ? mk(0.f) : DielEval(m, c);
        break;
    }
    case kMatSmoothCond:                                                              // :672-687
        o_wil = ReflectLocal(wol);
        o_pdf = kInf;
        o_comp = 1.f;
        o_att = mk(FresnelConductor(o_wil.z, m.eta, m.k));
        break;
    default: {                                                                        // :730-766
        const float F = FresnelDielectric(wol.z, m.eta);
        float a;
        if (u2 <= F) {
            o_wil = ReflectLocal(wol);
            a = F;
            o_comp = F;
        } else {
            bool inAbove;
            o_wil = Refract(wol, mk(0.f, 0.f, 1.f), m.eta, inAbove);
            a = (1.0f - F) * (inAbove ? Sqr(m.eta) : Sqr(m.etaInv));
            o_comp = 1.0f - F;
        }
        o_att = mk(a);
        o_pdf = kInf;
        break;
    }
    }
    r = BsdfRec{o_wil, o_att, o_pdf, o_comp};
}

PG3_HD float RRProb(const DevMaterial &m, F3 wol) {
    switch (m.type) {
    case kMatLambert:                                                                 // :301-322
        return Clampf(0.15f * Luminance(m.diffuse) + 0.85f * MaxC(m.diffuse), 0.f, 1.f);
    case kMatPhong: {                                                                 // :542-570
        const float d = 0.15f * Luminance(m.diffuse) + 0.85f * MaxC(m.diffuse);
        const float g = (0.15f * Luminance(m.glossy) + 0.85f * MaxC(m.glossy)) * (0.5f + 0.5f * fmaxf(wol.z, 0.f));
        return Clampf(d + g, 0.f, 1.f);
    }
    case kMatSmoothCond:                                                              // :691-700
        return FresnelConductor(wol.z, m.eta, m.k);
    default:
        return 1.0f;
    }
}
PG3_HD bool ReflectanceZero(const DevMaterial &m) {
    if (m.type == kMatLambert) return IsZero(m.diffuse);                              // :324-327
    if (m.type == kMatPhong) return IsZero(m.diffuse) && IsZero(m.glossy);           // :572-575
    return false;
}

}  // namespace pg3
