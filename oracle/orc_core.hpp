// Oracle (test infrastructure, never shipped or measured as the product): a CPU restatement of
// PG3Render's path-tracing hot path, used only by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg.  Each function cites the reference file:line whose behaviour it restates.
//
// This header: vector math, frames, the counter-based RNG, Fresnel, GGX microfacet math,
// sampling warps and the BSDFs.
//
// Parity unpinned: the random-number stream.  The reference draws from std::mt19937_64 through
// std::uniform_real_distribution<float> (rng.hxx:16-61), whose output the C++ standard leaves to
// the library; this restatement and the GPU both use Philox4x32-10 (checked against the Random123
// known-answer vectors) with the reference's draw order, so images agree with the reference's in
// distribution, not sample for sample.  Every other function here is checked against the
// reference's known answers or against closed forms (tests/test_oracle.py).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

#include "pg3g.h"

namespace orc {

const float kPiF = 3.14159265358979f;            // math.hxx:12
const float kPiInvF = 1.f / kPiF;                 // math.hxx:13
const float kPiDiv2F = 1.5707963f;                // math.hxx:14
const float kTinyF = 9.5367431640625E-7f;         // math.hxx:45 (2^-20)
const float kEpsRay = 1e-3f;                      // geom.hxx:16
const float kEpsDist = 1e-6f;                     // geom.hxx:27
const float kBlockerEps = 1e-5f;                  // materials.hxx:16
inline float InfF() { return std::numeric_limits<float>::infinity(); }
inline bool IsTiny(float x) { return std::abs(x) < kTinyF; }  // math.hxx:48-51
inline float Sqr(float x) { return x * x; }
inline float SafeSqrt(float x) { return std::sqrt(std::max(0.f, x)); }  // math.hxx:119-122
inline float Clamp(float x, float a, float b) { return std::min(std::max(x, a), b); }

struct V2 { float x, y; };

struct V3 {
    float x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(float a) : x(a), y(a), z(a) {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float &operator[](int i) { return (&x)[i]; }
    float operator[](int i) const { return (&x)[i]; }
    V3 operator+(const V3 &o) const { return V3(x + o.x, y + o.y, z + o.z); }
    V3 operator-(const V3 &o) const { return V3(x - o.x, y - o.y, z - o.z); }
    V3 operator*(const V3 &o) const { return V3(x * o.x, y * o.y, z * o.z); }
    V3 operator*(float s) const { return V3(x * s, y * s, z * s); }
    V3 operator/(float s) const { return V3(x / s, y / s, z / s); }
    V3 operator/(const V3 &o) const { return V3(x / o.x, y / o.y, z / o.z); }
    V3 operator-() const { return V3(-x, -y, -z); }
    V3 &operator+=(const V3 &o) { x += o.x; y += o.y; z += o.z; return *this; }
    V3 &operator*=(const V3 &o) { x *= o.x; y *= o.y; z *= o.z; return *this; }
    float Max() const { return std::max(std::max(x, y), z); }      // math.hxx:535
    bool IsZero() const { return x == 0 && y == 0 && z == 0; }     // math.hxx:536
    float LenSqr() const { return x * x + y * y + z * z; }
    float Length() const { return std::sqrt(LenSqr()); }
    // sRGB luminance, spectrum.hxx:74-80
    float Luminance() const { return 0.212671f * x + 0.715160f * y + 0.072169f * z; }
    // Vec3Base::ClipProportionally, math.hxx:538-549
    void ClipProportionally(float maxAllowed) {
        const float m = Max();
        if (m > maxAllowed) { const float s = maxAllowed / m; x *= s; y *= s; z *= s; }
    }
};
inline V3 operator*(float s, const V3 &v) { return v * s; }
inline float Dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 Cross(const V3 &a, const V3 &b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline V3 Normalize(const V3 &a) { return a / std::sqrt(Dot(a, a)); }  // math.hxx:653-658

// Frame, math.hxx:955-1015
struct Frame {
    V3 mX{1, 0, 0}, mY{0, 1, 0}, mZ{0, 0, 1};
    void SetFromZ(const V3 &z) {
        mZ = z;
        V3 tmpX = (std::abs(mZ.x) > 0.99f) ? V3(0, 1, 0) : V3(1, 0, 0);
        mY = Normalize(Cross(mZ, tmpX));
        mX = Cross(mY, mZ);
    }
    V3 ToWorld(const V3 &a) const { return mX * a.x + mY * a.y + mZ * a.z; }
    V3 ToLocal(const V3 &a) const { return V3(Dot(a, mX), Dot(a, mY), Dot(a, mZ)); }
    void SwitchNormal() { mZ = -mZ; mY = -mY; }
    const V3 &Normal() const { return mZ; }
};

// ---------------------------------------------------------------------------------------------
// Counter-based RNG.  The reference draws from one std::mt19937_64 stream per renderer
// (rng.hxx:16-61); its streams are implementation-defined, so parity is statistical and both
// this oracle and the GPU path draw from Philox4x32-10 (Salmon et al., SC'11) keyed by
// (seed) and counted by (pixel, iteration, vertex/block, scene|branch).  Each use of a random
// number owns a fixed dimension slot, so draws do not depend on evaluation order.
struct Philox {
    static inline void Round(uint32_t c[4], uint32_t k[2]) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k[0];
        uint32_t n2 = hi0 ^ c[3] ^ k[1];
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    }
    static inline void Gen(uint32_t out[4], uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                           uint64_t seed) {
        uint32_t c[4] = {c0, c1, c2, c3};
        uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
        for (int r = 0; r < 10; ++r) {
            Round(c, k);
            k[0] += 0x9E3779B9u;
            k[1] += 0xBB67AE85u;
        }
        out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
    }
};
inline float U32ToUnit(uint32_t v) { return (float)(v >> 8) * (1.0f / 16777216.0f); }

// Dimension slots of one path vertex.
enum Dim {
    kDimLightPick = 0, kDimLight0 = 1, kDimLight1 = 2, kDimLight2 = 3,
    kDimRR = 4, kDimBsdf0 = 5, kDimBsdf1 = 6, kDimBsdf2 = 7,
    kDimEstimateBase = 8,   // EnvironmentMap::EstimateIrradiance MC rounds (em.hxx:140-169)
};

struct PathRng {
    uint64_t seed;
    uint32_t pixel, iteration, scene, branch;
    // vertex: 0 = camera, k = path length k; sub: index of a split sample at the vertex
    float Get(uint32_t vertex, uint32_t sub, uint32_t dim) const {
        uint32_t out[4];
        const uint32_t c2 = (vertex << 16) | ((sub & 0xFFu) << 8) | (dim >> 2);
        const uint32_t c3 = (scene << 16) | (branch & 0xFFFFu);
        Philox::Gen(out, pixel, iteration, c2, c3, seed);
        return U32ToUnit(out[dim & 3]);
    }
};

// ---------------------------------------------------------------------------------------------
// Geometry helpers (geom.hxx)
inline float EpsRayCos(float c) { return (1.f + 2.f * (1.f - c)) * kEpsRay; }  // geom.hxx:23-26
inline V3 ReflectLocal(const V3 &d) { return V3(-d.x, -d.y, d.z); }            // geom.hxx:108-111
inline float TanThetaSqr(const V3 &d) {                                          // geom.hxx:96-106
    const float c2 = d.z * d.z;
    const float s2 = 1.f - c2;
    if (s2 <= 0.0f) return 0.0f;
    return s2 / c2;
}
// geom.hxx:119-133
inline void Reflect(V3 &out, bool &above, const V3 &in, const V3 &n) {
    const float dot = Dot(in, n);
    out = (2.0f * dot) * n - in;
    above = dot > 0.0f;
}
// geom.hxx:146-178
inline void Refract(V3 &out, bool &inAbove, const V3 &in, const V3 &n, float eta) {
    const float cosI = Dot(in, n);
    inAbove = (cosI > 0.0f);
    if (inAbove) eta = 1.0f / eta;
    const float cosT2 = 1 - (1 - cosI * cosI) * (eta * eta);
    if (cosT2 < 0.0f) { bool d; Reflect(out, d, in, n); return; }
    float cosT = std::sqrt(cosT2);
    cosT = (cosI > 0.0f) ? -cosT : cosT;
    out = n * (cosI * eta + cosT) - in * eta;
}
// geom.hxx:65-95
inline V3 CreateDirection(float sinT, float cosT, float sinP, float cosP) {
    return V3(sinT * cosP, sinT * sinP, cosT);
}
inline V3 LatLong2Dir(float u, float v) {
    const float phi = -(u - 0.5f) * 2 * kPiF;
    const float theta = v * kPiF;
    return CreateDirection(std::sin(theta), std::cos(theta), std::sin(phi), std::cos(phi));
}
inline V2 Dir2LatLong(const V3 &d) {
    const float phi = -std::atan2(d.y, d.x);
    const float theta = std::acos(d.z);
    const float u = Clamp(0.5f + phi * 0.5f * kPiInvF, 0.f, 1.f);
    const float v = Clamp(theta * kPiInvF, 0.f, 1.0f);
    return V2{u, v};
}

// ---------------------------------------------------------------------------------------------
// Sampling warps (sampling.hxx)
inline V3 SampleCosHemisphereW(float u0, float u1, float *pdf) {               // :289-307
    const float t1 = 2.f * kPiF * u0;
    const float t2 = std::sqrt(1.f - u1);
    V3 r(std::cos(t1) * t2, std::sin(t1) * t2, std::sqrt(u1));
    if (pdf) *pdf = r.z * kPiInvF;
    return r;
}
inline float CosHemispherePdfW(float z) { return std::max(0.f, z) * kPiInvF; }  // :309-312
inline V3 SampleCosSphereParamPdfW(float u0, float u1, float u2, bool up, bool low,
                                   float &pdf) {                                // :326-345
    V3 w = SampleCosHemisphereW(u0, u1, &pdf);
    if (up && low) { if (u2 < 0.5f) w = w * -1.0f; pdf *= 0.5f; }
    else if (low) w = w * -1.0f;
    return w;
}
inline float CosSpherePdfW(bool up, bool low, const V3 &d) {                    // :347-364
    if ((!up && d.z > 0.f) || (!low && d.z < 0.f)) return 0.f;
    float pdf = CosHemispherePdfW(std::abs(d.z));
    if (up && low) pdf *= 0.5f;
    return pdf;
}
inline V3 SamplePowerCosHemisphereW(float u0, float u1, float power) {         // :250-265
    const float t1 = 2.f * kPiF * u0;
    const float t2 = std::pow(u1, 1.f / (power + 1.f));
    const float t3 = std::sqrt(1.f - t2 * t2);
    return V3(std::cos(t1) * t3, std::sin(t1) * t3, t2);
}
inline float PowerCosHemispherePdfW(const V3 &d, float power) {                 // :276-284
    const float c = std::max(0.f, d.z);
    return (power + 1.f) * std::pow(c, power) * (kPiInvF * 0.5f);
}
// geom.hxx:814-827 (MapCartToBary) + sampling.hxx:66-75 (SampleUniformTriangle)
inline V3 SampleUniformTriangle(const V3 &p0, const V3 &p1, const V3 &p2, float u0, float u1) {
    const float xs = std::sqrt(u0);
    const float b0 = 1.f - xs, b1 = u1 * xs;
    const float b2 = 1.0f - b0 - b1;                                             // geom.hxx:687
    return p0 * b0 + p1 * b1 + p2 * b2;
}

// ---------------------------------------------------------------------------------------------
// Fresnel (physics.hxx)
inline float FresnelDielectric(float cosI, float etaAbs) {                      // :5-47
    if (etaAbs < 0) return 1.f;
    float we;
    if (cosI < 0.f) { cosI = -cosI; we = etaAbs; } else we = 1.f / etaAbs;
    const float sinT2 = Sqr(we) * (1.f - Sqr(cosI));
    const float cosT = SafeSqrt(1.f - sinT2);
    if (sinT2 > 1.0f) return 1.0f;
    const float t2 = we * cosI;
    const float rs = Sqr((t2 - cosT) / (t2 + cosT));
    const float t1 = we * cosT;
    const float rp = Sqr((cosI - t1) / (cosI + t1));
    return 0.5f * (rp + rs);
}
inline float FresnelConductor(float cosI, float eta, float k) {                 // :50-141 (Mitsuba form)
    if (cosI < -0.00001f) return 0.0f;
    cosI = Clamp(cosI, 0.0f, 1.0f);
    const float c2 = cosI * cosI, s2 = 1 - c2, s4 = s2 * s2;
    const float e2 = eta * eta, k2 = k * k;
    const float temp1 = e2 - k2 - s2;
    const float a2pb2 = SafeSqrt(temp1 * temp1 + 4 * k2 * e2);
    const float a = SafeSqrt(0.5f * (a2pb2 + temp1));
    const float term1 = a2pb2 + c2, term2 = 2 * a * cosI;
    const float Rs2 = (term1 - term2) / (term1 + term2);
    const float term3 = a2pb2 * c2 + s4, term4 = term2 * s2;
    const float Rp2 = Rs2 * (term3 - term4) / (term3 + term4);
    return 0.5f * (Rp2 + Rs2);
}

// ---------------------------------------------------------------------------------------------
// GGX microfacet math (microfacet.hxx)
inline float ReflectionJacobian(const V3 &wo, const V3 &h) {                    // :16-31
    const float c = std::max(Dot(h, wo), 0.000001f);
    return 1.0f / (4.0f * c);
}
inline float RefractionJacobian(const V3 &wi, const V3 &wo, const V3 &h, float etaOutIn) {  // :34-56
    const float cOM = Dot(h, wo), cIM = Dot(h, wi);
    const float den = Sqr(etaOutIn * cIM + cOM);
    return std::abs(cOM) / std::max(den, 0.000001f);
}
inline V3 HalfwayReflection(const V3 &wi, const V3 &wo) {                       // :60-84
    V3 h = wi + wo;
    const float len = h.Length();
    if (IsTiny(len)) h = V3(0.0f, 0.0f, 1.0f); else h = h / len;
    return h * ((h.z >= 0.0f) ? 1.0f : -1.0f);
}
inline V3 HalfwayRefraction(const V3 &wi, const V3 &wo, float etaInOut) {       // :88-123
    V3 h = wi + wo * etaInOut;
    const float len = h.Length();
    if (IsTiny(len)) h = wi; else h = h / len;
    return h * ((h.z >= 0.0f) ? 1.0f : -1.0f);
}
inline float DistributionGgx(const V3 &m, float alpha) {                        // :505-527
    if (m.z <= 0.f) return 0.0f;
    const float a2 = alpha * alpha;
    const float c2 = m.z * m.z;
    const float t2 = TanThetaSqr(m);
    const float t1 = a2 + t2;
    const float t = c2 * t1;
    return a2 / (kPiF * t * t);
}
inline float SmithG1(const V3 &v, const V3 &m, float alpha) {                   // :529-554
    if (m.z <= 0) return 0.0f;
    const float cVM = Dot(v, m);
    if ((v.z * cVM) < 0.0f) return 0.0f;
    const float a2 = alpha * alpha;
    const float t2 = TanThetaSqr(v);
    const float root = std::sqrt(1.0f + a2 * t2);
    return 2.0f / (1.0f + root);
}
inline void SampleGgxP11(V2 &slope, float thetaI, float ux, float uy) {        // :605-669
    if (thetaI < 0.0001f) {
        const float sx = std::min(ux, 0.9999f);
        const float r = SafeSqrt(sx / (1 - sx));
        const float phi = 2 * kPiF * uy;
        slope = V2{r * std::cos(phi), r * std::sin(phi)};
        return;
    }
    const float tanI = std::tan(thetaI);
    const float tanIInv = 1.0f / tanI;
    const float G1 = 2.0f / (1.0f + SafeSqrt(1.0f + 1.0f / (tanIInv * tanIInv)));
    float A = 2.0f * ux / G1 - 1.0f;
    if (std::abs(A) == 1.0f) A -= (std::signbit(A) ? -1.0f : 1.0f) * 1e-4f;
    const float B = tanI;
    const float tmp = 1.0f / (A * A - 1.0f);
    const float D = SafeSqrt(B * B * tmp * tmp - (A * A - B * B) * tmp);
    const float sx1 = B * tmp - D, sx2 = B * tmp + D;
    slope.x = (A < 0.0f || sx2 > (1.0f / tanI)) ? sx1 : sx2;
    float ys, yh;
    if (uy > 0.5f) { ys = 1.0f; yh = 2.0f * (uy - 0.5f); }
    else { ys = -1.0f; yh = 2.0f * (0.5f - uy); }
    // rational fit of the inverse conditional CDF (Mitsuba), coefficients as in microfacet.hxx:655-662
    const float z =
        (yh * (yh * (yh * -(float)0.365728915865723 + (float)0.790235037209296) -
               (float)0.424965825137544) + (float)0.000152998850436920) /
        (yh * (yh * (yh * (yh * (float)0.169507819808272 - (float)0.397203533833404) -
                     (float)0.232500544458471) + (float)1) - (float)0.539825872510702);
    slope.y = ys * z * std::sqrt(1.0f + slope.x * slope.x);
}
inline V3 SampleGgxVisibleNormals(const V3 &wo, float alpha, float ux, float uy) {  // :678-729
    const V3 ws = Normalize(V3(wo.x * alpha, wo.y * alpha, std::max(wo.z, 0.0f)));
    float theta = 0.0f, phi = 0.0f;
    if (ws.z < 0.999f) { theta = std::acos(ws.z); phi = std::atan2(ws.y, ws.x); }
    const float sp = std::sin(phi), cp = std::cos(phi);
    V2 s;
    SampleGgxP11(s, theta, ux, uy);
    const V2 r{s.x * cp - s.y * sp, s.x * sp + s.y * cp};
    const V2 sl{r.x * alpha, r.y * alpha};
    const float inv = 1.0f / V3(sl.x, sl.y, 1.0f).Length();
    return V3(-sl.x * inv, -sl.y * inv, inv);
}
inline float GgxVndfPdf(const V3 &wo, const V3 &h, float D, float alpha) {      // :732-756
    if (h.z <= 0.f) return 0.0f;
    const float G1 = SmithG1(wo, h, alpha);
    const float cOM = Dot(h, wo);
    const float cO = std::max(wo.z, 0.00001f);
    return (G1 * std::abs(cOM) * D) / cO;
}

// ---------------------------------------------------------------------------------------------
// Materials (materials.hxx)
struct MatRec {        // MaterialRecord, materials.hxx:48-207
    V3 wol, wil, att;
    float pdfW = 0.f, compProb = 0.f;
    bool IsBlocker() const { return att.Max() * std::abs(wil.z) <= kBlockerEps; }
    float CosAbs() const { return std::abs(wil.z); }
    bool IsFinite() const { return pdfW != InfF(); }
};

struct Material {
    int type = PG3G_MAT_LAMBERT;
    V3 diffuse, glossy{1.f}, tint{1.f};
    float phongExp = 1.f;
    float alpha = 0.1f, eta = 1.f, etaInv = 1.f, k = 0.f;
    bool front = true, back = false;   // MaterialProperties light-sampling sides
    int lightId = -1;

    static Material FromDesc(const pg3g_material &d);
    // EvalBsdf with kOptSamplingProbs requested
    void Eval(MatRec &r) const;
    void Sample(float u0, float u1, float u2, MatRec &r) const;
    float RRProb(const V3 &wol) const;
    bool ReflectanceZero() const;
};

inline Material Material::FromDesc(const pg3g_material &d) {
    Material m;
    m.type = d.type;
    m.lightId = d.light_id;
    const float outer = d.outer_ior, inner = d.inner_ior;
    switch (d.type) {
    case PG3G_MAT_LAMBERT:                                                   // :253-266
        m.diffuse = V3(d.diffuse[0], d.diffuse[1], d.diffuse[2]);
        m.front = true; m.back = false;
        break;
    case PG3G_MAT_PHONG:                                                     // :364-379
        m.diffuse = d.phong_diffuse_on ? V3(d.diffuse[0], d.diffuse[1], d.diffuse[2]) : V3(0.f);
        m.glossy = d.phong_glossy_on ? V3(d.glossy[0], d.glossy[1], d.glossy[2]) : V3(0.f);
        m.phongExp = d.phong_exponent;
        if (d.phong_glossy_on) m.diffuse = m.diffuse / 2;
        m.front = true; m.back = false;
        break;
    case PG3G_MAT_GGX_CONDUCTOR:                                             // :791-806
        m.eta = !IsTiny(outer) ? inner / outer : 1.0f;
        m.k = d.absorbance;
        m.alpha = Clamp(d.alpha, 0.001f, 1.0f);
        m.tint = V3(d.glossy[0], d.glossy[1], d.glossy[2]);
        m.front = true; m.back = false;
        break;
    case PG3G_MAT_GGX_DIELECTRIC:                                            // :996-1021
        if (!IsTiny(outer)) { m.eta = inner / outer; m.etaInv = outer / inner; }
        else { m.eta = 1.0f; m.etaInv = 1.0f; }
        m.alpha = Clamp(d.alpha, 0.001f, 1.0f);
        m.front = true; m.back = d.back_side_light_sampling != 0;
        break;
    case PG3G_MAT_SMOOTH_CONDUCTOR:                                          // :660-669
        m.eta = !IsTiny(outer) ? inner / outer : 1.0f;
        m.k = d.absorbance;
        m.front = false; m.back = false;
        break;
    case PG3G_MAT_SMOOTH_DIELECTRIC:                                         // :716-727
        m.eta = !IsTiny(outer) ? inner / outer : 1.0f;
        m.etaInv = !IsTiny(inner) ? outer / inner : 1.0f;
        m.front = false; m.back = false;
        break;
    }
    return m;
}

// Phong helpers, materials.hxx:381-409 and :579-596
inline float PhongPdf(const Material &m, const V3 &wol, const V3 &wil, float dEst, float gEst) {
    const float tot = dEst + gEst;
    if (tot < kBlockerEps) return CosHemispherePdfW(wil.z);
    const V3 wrl = ReflectLocal(wol);
    Frame lf; lf.SetFromZ(wrl);
    const V3 wc = lf.ToLocal(wil);
    const float dp = dEst / tot, gp = gEst / tot;
    return dp * CosHemispherePdfW(wil.z) + gp * PowerCosHemispherePdfW(wc, m.phongExp);
}
inline V3 PhongEval(const Material &m, const V3 &wil, const V3 &wol) {
    if (wil.z <= 0.f || wol.z <= 0.f) return V3(0.f);
    const V3 d = m.diffuse / kPiF;
    const float cc = (m.phongExp + 2.0f) / (2.0f * kPiF);
    const V3 wrl = ReflectLocal(wil);
    const float c = std::max(Dot(wrl, wol), 0.f);
    return d + m.glossy * (cc * std::pow(c, m.phongExp));
}

// GGX conductor eval, materials.hxx:889-938 (with an RGB tint that is (1,1,1) in the reference)
inline V3 GgxCondEvalCtx(const Material &m, const V3 &wil, const V3 &wol, const V3 &h, float D,
                         float F) {
    if (wil.z <= 0.f || wol.z <= 0.f || IsTiny(4.0f * wil.z * wol.z)) return V3(0.f);
    const float G = SmithG1(wil, h, m.alpha) * SmithG1(wol, h, m.alpha);
    const float v = (F * G * D) / (4.0f * wil.z * wol.z);
    return m.tint * v;
}

struct DielCtx {      // materials.hxx:1118-1132
    V3 wolS, wilS, h;
    bool below = false, refl = true;
    float etaS = 1, etaInvS = 1, F = 0, D = 0;
};
inline V3 DielEval(const Material &m, const DielCtx &c) {                        // :1169-1225
    const float ci = std::abs(c.wilS.z), co = std::abs(c.wolS.z);
    if ((!c.refl && IsTiny(ci * co)) || (c.refl && IsTiny(4.0f * ci * co))) return V3(0.f);
    const float G = SmithG1(c.wilS, c.h, m.alpha) * SmithG1(c.wolS, c.h, m.alpha);
    float v;
    if (c.refl) {
        v = (c.F * G * c.D) / (4.0f * ci * co);
    } else {
        const float cMI = Dot(c.h, c.wilS), cMO = Dot(c.h, c.wolS);
        v = ((std::abs(cMI) * std::abs(cMO)) / (ci * co)) *
            ((Sqr(c.etaInvS) * (1.0f - c.F) * G * c.D) / (Sqr(cMI + c.etaInvS * cMO)));
        v *= Sqr(c.etaS);
    }
    return V3(v);
}
inline float DielPdf(const Material &m, const DielCtx &c) {                      // :1227-1266
    if (c.wolS.z < 0.f) return 0.0f;
    const float vn = GgxVndfPdf(c.wolS, c.h, c.D, m.alpha);
    const float jac = c.refl ? ReflectionJacobian(c.wilS, c.h)
                             : RefractionJacobian(c.wolS, c.wilS, c.h, c.etaInvS);
    const float cp = c.refl ? c.F : (1.f - c.F);
    return vn * jac * cp;
}

inline void Material::Eval(MatRec &r) const {
    switch (type) {
    case PG3G_MAT_LAMBERT:                                                   // :268-345
        r.att = (r.wil.z <= 0.f || r.wol.z <= 0.f) ? V3(0.f) : diffuse / kPiF;
        r.pdfW = CosHemispherePdfW(r.wil.z);
        r.compProb = 1.f;
        break;
    case PG3G_MAT_PHONG: {                                                   // :424-437
        r.att = PhongEval(*this, r.wil, r.wol);
        const float dEst = diffuse.Luminance();
        const float cto = std::max(r.wol.z, 0.f);
        const float gEst = glossy.Luminance() * (0.5f + 0.5f * cto);
        r.pdfW = PhongPdf(*this, r.wol, r.wil, dEst, gEst);
        r.compProb = 1.0f;
        break;
    }
    case PG3G_MAT_GGX_CONDUCTOR: {                                           // :808-816, 889-959
        const V3 h = HalfwayReflection(r.wil, r.wol);
        const float D = DistributionGgx(h, alpha);
        const float F = FresnelConductor(Dot(h, r.wol), eta, k);
        r.att = GgxCondEvalCtx(*this, r.wil, r.wol, h, D, F);
        r.compProb = 1.0f;
        if (r.wol.z < 0.f) r.pdfW = 0.0f;
        else r.pdfW = GgxVndfPdf(r.wol, h, D, alpha) * ReflectionJacobian(r.wol, h);
        break;
    }
    case PG3G_MAT_GGX_DIELECTRIC: {                                          // :1023-1031, 1134-1167
        DielCtx c;
        c.below = (r.wol.z < 0.f);
        c.refl = ((r.wil.z > 0.f) == (r.wol.z > 0.f));
        c.wolS = r.wol * (c.below ? -1.0f : 1.0f);
        c.wilS = r.wil * (c.below ? -1.0f : 1.0f);
        c.etaS = c.below ? etaInv : eta;
        c.etaInvS = c.below ? eta : etaInv;
        c.h = c.refl ? HalfwayReflection(c.wilS, c.wolS) : HalfwayRefraction(c.wilS, c.wolS, c.etaInvS);
        c.D = DistributionGgx(c.h, alpha);
        c.F = FresnelDielectric(Dot(c.h, c.wolS), c.etaS);
        r.att = DielEval(*this, c);
        r.pdfW = DielPdf(*this, c);
        r.compProb = 1.0f;
        break;
    }
    default:                                                                 // :612-654
        r.att = V3(0.f);
        r.pdfW = 0.0f;
        r.compProb = 0.0f;
        break;
    }
}

inline void Material::Sample(float u0, float u1, float u2, MatRec &r) const {
    switch (type) {
    case PG3G_MAT_LAMBERT:                                                   // :285-292
        r.wil = SampleCosHemisphereW(u0, u1, &r.pdfW);
        r.att = (r.wil.z <= 0.f || r.wol.z <= 0.f) ? V3(0.f) : diffuse / kPiF;
        r.compProb = 1.f;
        break;
    case PG3G_MAT_PHONG: {                                                   // :442-504
        const float dEst = diffuse.Luminance();
        const float cto = std::max(r.wol.z, 0.f);
        const float gEst = glossy.Luminance() * (0.5f + 0.5f * cto);
        const float tot = dEst + gEst;
        if (tot < kBlockerEps) {
            r.att = V3(0.f);
            r.wil = SampleCosHemisphereW(u0, u1, &r.pdfW);
            r.compProb = 1.f;
            return;
        }
        const float rv = u2 * tot;
        if (rv < dEst) {
            r.wil = SampleCosHemisphereW(u0, u1, nullptr);
        } else {
            const V3 cs = SamplePowerCosHemisphereW(u0, u1, phongExp);
            const V3 wrl = ReflectLocal(r.wol);
            Frame lf; lf.SetFromZ(wrl);
            r.wil = lf.ToWorld(cs);
        }
        r.compProb = 1.f;
        r.pdfW = PhongPdf(*this, r.wol, r.wil, dEst, gEst);
        r.att = (r.wil.z > 0.0f) ? PhongEval(*this, r.wil, r.wol) : V3(0.f);
        break;
    }
    case PG3G_MAT_GGX_CONDUCTOR: {                                           // :819-849
        const V3 h = SampleGgxVisibleNormals(r.wol, alpha, u0, u1);
        const float D = DistributionGgx(h, alpha);
        bool above;
        V3 wil;
        Reflect(wil, above, r.wol, h);
        r.wil = wil;
        const float F = FresnelConductor(Dot(h, r.wol), eta, k);
        r.compProb = 1.0f;
        r.pdfW = (r.wol.z < 0.f) ? 0.0f : GgxVndfPdf(r.wol, h, D, alpha) * ReflectionJacobian(r.wol, h);
        if (!above || wil.z < 0.0f) r.att = V3(0.f);
        else r.att = GgxCondEvalCtx(*this, wil, r.wol, h, D, F);
        break;
    }
    case PG3G_MAT_GGX_DIELECTRIC: {                                          // :1034-1089
        DielCtx c;
        c.below = (r.wol.z < 0.f);
        c.wolS = r.wol * (c.below ? -1.0f : 1.0f);
        c.etaS = c.below ? etaInv : eta;
        c.etaInvS = c.below ? eta : etaInv;
        c.h = SampleGgxVisibleNormals(c.wolS, alpha, u0, u1);
        c.D = DistributionGgx(c.h, alpha);
        c.F = FresnelDielectric(Dot(c.h, c.wolS), c.etaS);
        bool above;
        float cosAbs;
        if (u2 <= c.F) {
            Reflect(c.wilS, above, c.wolS, c.h);
            cosAbs = c.wilS.z;
            c.refl = true;
        } else {
            Refract(c.wilS, above, c.wolS, c.h, c.etaS);
            cosAbs = -c.wilS.z;
            c.refl = false;
        }
        r.wil = c.wilS * (c.below ? -1.0f : 1.0f);
        r.pdfW = DielPdf(*this, c);
        r.compProb = 1.0f;
        if (!above || cosAbs < 0.0f) r.att = V3(0.f);
        else r.att = DielEval(*this, c);
        break;
    }
    case PG3G_MAT_SMOOTH_CONDUCTOR: {                                        // :672-687
        r.wil = ReflectLocal(r.wol);
        r.pdfW = InfF();
        r.compProb = 1.f;
        r.att = V3(FresnelConductor(r.wil.z, eta, k));
        break;
    }
    case PG3G_MAT_SMOOTH_DIELECTRIC: {                                       // :730-766
        const float F = FresnelDielectric(r.wol.z, eta);
        float a;
        if (u2 <= F) {
            r.wil = ReflectLocal(r.wol);
            a = F;
            r.compProb = F;
        } else {
            bool inAbove;
            Refract(r.wil, inAbove, r.wol, V3(0.f, 0.f, 1.f), eta);
            a = 1.0f - F;
            a *= inAbove ? Sqr(eta) : Sqr(etaInv);
            r.compProb = 1.0f - F;
        }
        r.att = V3(a);
        r.pdfW = InfF();
        break;
    }
    }
}

inline float Material::RRProb(const V3 &wol) const {
    switch (type) {
    case PG3G_MAT_LAMBERT: {                                                 // :301-322
        const float p = 0.15f * diffuse.Luminance() + 0.85f * diffuse.Max();
        return Clamp(p, 0.f, 1.f);
    }
    case PG3G_MAT_PHONG: {                                                   // :542-570
        const float dEst = 0.15f * diffuse.Luminance() + 0.85f * diffuse.Max();
        const float cto = std::max(wol.z, 0.f);
        const float gEst = (0.15f * glossy.Luminance() + 0.85f * glossy.Max()) * (0.5f + 0.5f * cto);
        return Clamp(dEst + gEst, 0.f, 1.f);
    }
    case PG3G_MAT_SMOOTH_CONDUCTOR:                                          // :691-700
        return FresnelConductor(wol.z, eta, k);
    default:
        return 1.0f;
    }
}

inline bool Material::ReflectanceZero() const {
    switch (type) {
    case PG3G_MAT_LAMBERT: return diffuse.IsZero();                          // :324-327
    case PG3G_MAT_PHONG: return diffuse.IsZero() && glossy.IsZero();         // :572-575
    default: return false;
    }
}

}  // namespace orc
