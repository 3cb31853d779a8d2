// Device scene layer: flat SoA-friendly scene records in HBM, brute-force intersection,
// lights, the environment-map image and its two table-driven samplers (2-D luminance CDF and the
// steerable 6-ary triangle tree).  Behaviour follows the cited PG3Render files.
#pragma once

#include "dev_math.hpp"

namespace pg3 {

// Triangle record: 96 B, read with wave-uniform (scalar) loads in the brute-force loop.  The
// edges are stored as Pluecker lines: the reference's edge test (scene_graph.hxx:107-120),
// dot(cross(pj - o, pi - o), d), equals dot(d, pj x pi) - dot(o x d, pj - pi), so with the ray's
// moment o x d computed once per ray each edge costs two 3-term dot products instead of three
// subtractions, a cross product and a dot product.  Edge k: u[k] = pj - pi, v[k] = pj x pi for
// (j, i) = (2, 1), (1, 0), (0, 2).  np0 = dot(n, p0) gives the plane distance.
struct alignas(16) DevTri {
    float u[3][3], v[3][3];
    float n[3];          // normalize(cross(p1-p0, p2-p0)), scene_graph.hxx:100
    float np0;
    int32_t mat;
    int32_t pad[3];
};
// Axis-aligned rectangle made of two triangles of the scene (the walls of the reference's boxes,
// scene.hxx:459-488): both triangles lie in the plane x_A = const, have bit-identical normals
// (exactly zero off axis A) and plane distances, the same material, and together cover
// [lo0, hi0] x [lo1, hi1] along the axes (A + 1) % 3 and (A + 2) % 3.  The plane distance of a hit
// takes the triangles' own float operations, so hit distance, normal and material are the
// triangles' bit for bit; one bounds test replaces the two triangles' edge tests, which differ from
// it only for rays within a few ulp of the rectangle's border.  48 B, read with scalar loads.
struct alignas(16) DevQuad {
    float n[3];
    float np0;
    float lo0, hi0, lo1, hi1;
    int32_t mat;
    int32_t pad[3];
};
struct alignas(16) DevSphere {
    float c[3], r;
    int32_t mat;
    int32_t pad[3];
};
enum : int32_t { kLightArea = 0, kLightPoint = 1, kLightBackground = 2 };
struct alignas(16) DevLight {
    int32_t type;
    float invArea;
    float pad[2];
    F3 p0, e1, e2;       // area triangle (point light: position in p0)
    F3 n;                // AreaLight frame normal
    F3 radiance;         // area: Le = power*invArea/pi; point: intensity; background: constant radiance
};

enum : int32_t { kEmAbsent = -1, kEmNone = 0, kEmCosine = 1, kEmSimple = 2, kEmSteerable = 3 };
constexpr int kMaxTreeLevels = 16;

// Environment map and its sampling tables.
struct DevEm {
    uint32_t w, h;
    int32_t bilinear;
    int32_t sampler;
    const float4 *texels;      // w*h RGB texels padded to 16 B (one load per texel)
    const float *lum;          // w*h texel luminances (the steerable pdf reads luminance only)
    // simple spherical sampler (em_simple_spherical_sampler.hxx): FP32 CDFs built from FP64 sums
    const float *margCdf;      // h+1
    const float *condCdf;      // h*(w+1)
    float plan2ang;
    // steerable sampler (em_steerable_sampler.hxx): layered 6-ary tree
    uint32_t numLevels;        // levels[0] = leaves ... levels[numLevels-1] = root
    uint32_t levelBase[kMaxTreeLevels];
    uint32_t levelCount[kMaxTreeLevels];
    const float *nodeW;        // 12 floats per node (9 SH weights + pad), all levels concatenated
    const float *leafVerts;    // per leaf: 3 vertices x (dir 3 + SH weight 9) = 36 floats
    float rootW[9];
};

// Material classes of the wavefront pipeline's shading stage, one kernel each: diffuse (Lambert,
// Phong), GGX conductor, GGX dielectric, Dirac (smooth conductor and dielectric), and escaped
// paths.  kClsAny: a kernel that shades every class.
enum : int { kClsAny = -1, kClsDiffuse = 0, kClsGgxCond = 1, kClsGgxDiel = 2, kClsDirac = 3, kClsMiss = 4, kNumCls = 5 };
__host__ __device__ constexpr uint32_t ClassOfType(int32_t t) {
    return (t == kMatLambert || t == kMatPhong) ? kClsDiffuse
           : (t == kMatGgxCond)                 ? kClsGgxCond
           : (t == kMatGgxDiel)                 ? kClsGgxDiel
                                                : kClsDirac;
}

struct SceneView {
    const DevQuad *quads;      // rectangles normal to x, then y, then z: [0, quadEnd[0]), ...
    uint32_t quadEnd[3];
    const DevTri *tris;        // the triangles not merged into rectangles, in scene order
    const DevSphere *sphs;
    const DevMaterial *mats;
    const DevLight *lights;
    uint32_t numTris, numSphs, numLights;
    uint32_t matClasses;       // bit c: some material of class c (ClassOfType); kClsMiss always set
    int32_t bgLight;
    DevEm em;
    float r2w[16];             // raster->world (row-major), camera.hxx:47-55
    F3 camPos;
    uint32_t resX, resY;
    // Config
    int32_t algorithm;
    uint32_t minPL, maxPL;
    float iic;
    float splitBudget, splitLevel, lightRatio;   // -sb, splitting level, light-to-BSDF sample ratio
    uint32_t k0, k1, sceneId;
    // Differentiable rendering with frozen sampling: integrand values (BSDF values of Lambert and
    // GGX conductor materials, EM radiance) come from evalMats / evalTexels, while directions,
    // pdfs and MIS weights come from mats and the EM sampling tables.  Both pairs are equal unless
    // parameters were changed without refreshing the sampling tables.  gradBase[m] is the first
    // scalar gradient of material m (-1: not differentiated).
    const DevMaterial *evalMats;
    const float4 *evalTexels;
    const int32_t *gradBase;
};

// ---------------------------------------------------------------------------------------------
// Intersection (scene_graph.hxx, scene.hxx)
struct Hit {
    float dist;
    int32_t mat;
    F3 n;
};

// The intersection routines are built with FMA contraction and a fast reciprocal division (the
// rest of the device code is not): they are the bulk of the arithmetic, and a few-ulp hit distance
// is invisible in images; the function-level checks (pg3g_intersect_batch / pg3g_occluded_batch
// against the oracle) hold hit/miss and material exactly and distances to 1e-5 relative.
// m = org x dir, the ray's moment (computed once per ray).
PG3_DEV bool TriHit(const DevTri &t, F3 org, F3 dir, F3 m, float tmin, Hit &h) {  // :103-134
#pragma clang fp contract(fast)
    const float v0d = Dot(dir, mk(t.v[0][0], t.v[0][1], t.v[0][2])) - Dot(m, mk(t.u[0][0], t.u[0][1], t.u[0][2]));
    const float v1d = Dot(dir, mk(t.v[1][0], t.v[1][1], t.v[1][2])) - Dot(m, mk(t.u[1][0], t.u[1][1], t.u[1][2]));
    const float v2d = Dot(dir, mk(t.v[2][0], t.v[2][1], t.v[2][2])) - Dot(m, mk(t.u[2][0], t.u[2][1], t.u[2][2]));
    if (((v0d < 0.f) && (v1d < 0.f) && (v2d < 0.f)) || ((v0d >= 0.f) && (v1d >= 0.f) && (v2d >= 0.f))) {
        const F3 n = mk(t.n[0], t.n[1], t.n[2]);
        const float d = FastDiv(t.np0 - Dot(n, org), Dot(n, dir));
        if ((d > tmin) & (d < h.dist)) { h.dist = d; h.mat = t.mat; h.n = n; return true; }
    }
    return false;
}
// FP64 quotient and square root from the hardware estimates (v_rcp_f64, v_rsq_f64), each refined
// by two Newton steps and one residual correction: within a few ulp of the correctly rounded FP64
// result, so a root rounded to float agrees with the correctly rounded one unless it lies within
// ~1e-15 relative of a float rounding boundary.  About half the FP64 instructions of the
// correctly rounded sequences (scale, two Newton steps, fused correction, fix-up).
PG3_DEV double DivD(double a, double b) {
    if (!(fabs(b) > 0.0 && fabs(b) < 1e300)) return a / b;   // 0, inf, NaN, huge: exact
    double r = __builtin_amdgcn_rcp(b);
    r = fma(r, fma(-b, r, 1.0), r);
    r = fma(r, fma(-b, r, 1.0), r);
    const double q = a * r;
    return fma(r, fma(-b, q, a), q);
}
PG3_DEV double SqrtD(double x) {
    if (!(x > 0.0 && x < 1e300)) return sqrt(x);   // 0, negative, inf, NaN: exact
    double y = __builtin_amdgcn_rsq(x);
    y = y * fma(-0.5 * x * y, y, 1.5);
    y = y * fma(-0.5 * x * y, y, 1.5);
    const double s = x * y;
    return fma(0.5 * y, fma(-s, s, x), s);
}
// :176-245.  B*B - 4*A*C is evaluated in FP32 (all float operands) and widened; the root, q and
// both roots are FP64 as in the reference.
PG3_DEV bool SphHit(const DevSphere &s, F3 org, F3 dir, float tmin, Hit &h) {
#pragma clang fp contract(fast)
    const F3 o = org - mk(s.c[0], s.c[1], s.c[2]);
    const float A = Dot(dir, dir);
    const float B = 2 * Dot(dir, o);
    const float C = Dot(o, o) - (s.r * s.r);
    const float discF = B * B - 4 * A * C;
    if (discF < 0.f) return false;
    const double ds = SqrtD((double)discF);
    const double q = (B >= 0) ? ((-(double)B - ds) * 0.5) : ((-(double)B + ds) * 0.5);
    double t0 = DivD(q, (double)A);
    double t1 = (fabs(q) > 0.1) ? DivD((double)C, q) : t0;
    if (t0 > t1) { const double t = t0; t0 = t1; t1 = t; }
    float resT;
    if (t0 > tmin && t0 < h.dist) resT = (float)t0;
    else if (t1 > tmin && t1 < h.dist) resT = (float)t1;
    else return false;
    h.dist = resT;
    h.mat = s.mat;
    const F3 nn = o + mk(resT) * dir;
    h.n = FastDiv(nn, sqrtf(Dot(nn, nn)));
    return true;
}
// Primitive records are read through the constant address space: every lane of a wavefront
// tests the same primitive in the same order, so the loads are wave-uniform and become scalar
// loads into SGPRs (the tables are never written while a kernel runs).
template <class T> using ConstPtr = const __attribute__((address_space(4))) T *;
template <class T> PG3_DEV ConstPtr<T> AsConst(const T *p) { return (ConstPtr<T>)(p); }
// Scene tables reached through pointers stored in memory (the batched kernel reads its scenes
// from an array) are generic pointers to the compiler; casting them to the global address space
// gives global loads instead of flat loads.
template <class T> using GlobalPtr = const __attribute__((address_space(1))) T *;
template <class T> PG3_DEV GlobalPtr<T> AsGlobal(const T *p) { return (GlobalPtr<T>)(p); }
template <class T> PG3_DEV T LoadGlobal(const T *p) {
    static_assert(sizeof(T) % 4 == 0, "4-byte records only");
    const GlobalPtr<uint32_t> src = (GlobalPtr<uint32_t>)(p);
    T out;
    uint32_t *dst = reinterpret_cast<uint32_t *>(&out);
#pragma unroll
    for (size_t i = 0; i < sizeof(T) / 4; ++i) dst[i] = src[i];
    return out;
}
typedef float F4v __attribute__((ext_vector_type(4)));
PG3_DEV DevTri LoadTri(const DevTri *base, uint32_t i) {
    const ConstPtr<F4v> q = (ConstPtr<F4v>)AsConst(base + i);
    float f[24];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const F4v a = q[k];
        f[4 * k] = a.x; f[4 * k + 1] = a.y; f[4 * k + 2] = a.z; f[4 * k + 3] = a.w;
    }
    DevTri t;
#pragma unroll
    for (int e = 0; e < 3; ++e)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            t.u[e][c] = f[3 * e + c];
            t.v[e][c] = f[9 + 3 * e + c];
        }
    t.n[0] = f[18]; t.n[1] = f[19]; t.n[2] = f[20];
    t.np0 = f[21];
    t.mat = __float_as_int(f[22]);
    return t;
}
PG3_DEV DevQuad LoadQuad(const DevQuad *base, uint32_t i) {
    const ConstPtr<F4v> q = (ConstPtr<F4v>)AsConst(base + i);
    const F4v a = q[0], b = q[1], c = q[2];
    DevQuad r;
    r.n[0] = a.x; r.n[1] = a.y; r.n[2] = a.z; r.np0 = a.w;
    r.lo0 = b.x; r.hi0 = b.y; r.lo1 = b.z; r.hi1 = b.w;
    r.mat = __float_as_int(c.x);
    return r;
}
template <int A> PG3_DEV float Comp(F3 v) { return A == 0 ? v.x : A == 1 ? v.y : v.z; }
// Hit of the rectangle normal to axis A.  The distance is FastDiv(np0 - Dot(n, org), Dot(n, dir))
// as in TriHit: with n zero off axis A each dot product is the single rounded product of the A
// components (a product with zero adds nothing), so it is computed as that product, without
// contraction (np0 - n_A o_A is rounded after the product, as the dot product is).
template <int A>
PG3_DEV bool QuadHit(const DevQuad &q, F3 org, F3 dir, float tmin, Hit &h) {
    const float na = q.n[A];
    const float d = FastDiv(q.np0 - na * Comp<A>(org), na * Comp<A>(dir));
    if ((d > tmin) & (d < h.dist)) {
        const float pu = Comp<(A + 1) % 3>(org) + d * Comp<(A + 1) % 3>(dir);
        const float pv = Comp<(A + 2) % 3>(org) + d * Comp<(A + 2) % 3>(dir);
        if ((pu >= q.lo0) & (pu <= q.hi0) & (pv >= q.lo1) & (pv <= q.hi1)) {
            h.dist = d;
            h.mat = q.mat;
            h.n = mk(q.n[0], q.n[1], q.n[2]);
            return true;
        }
    }
    return false;
}
PG3_DEV DevSphere LoadSphere(const DevSphere *base, uint32_t i) {
    const ConstPtr<F4v> q = (ConstPtr<F4v>)AsConst(base + i);
    const F4v a = q[0], b = q[1];
    DevSphere s;
    s.c[0] = a.x; s.c[1] = a.y; s.c[2] = a.z; s.r = a.w;
    s.mat = __float_as_int(b.x);
    return s;
}

PG3_DEV F3 RayMoment(F3 org, F3 dir) {
#pragma clang fp contract(fast)
    return Cross(org, dir);
}
PG3_DEV bool Intersect(const SceneView &S, F3 org, F3 dir, float tmin, Hit &h) {  // scene.hxx:42-59
    bool any = false;
    for (uint32_t i = 0; i < S.quadEnd[2]; ++i) {
        const DevQuad q = LoadQuad(S.quads, i);
        if (i < S.quadEnd[0]) any |= QuadHit<0>(q, org, dir, tmin, h);
        else if (i < S.quadEnd[1]) any |= QuadHit<1>(q, org, dir, tmin, h);
        else any |= QuadHit<2>(q, org, dir, tmin, h);
    }
    const F3 m = RayMoment(org, dir);
    for (uint32_t i = 0; i < S.numTris; ++i) {
        const DevTri t = LoadTri(S.tris, i);
        any |= TriHit(t, org, dir, m, tmin, h);
    }
    for (uint32_t i = 0; i < S.numSphs; ++i) {
        const DevSphere s = LoadSphere(S.sphs, i);
        any |= SphHit(s, org, dir, tmin, h);
    }
    return any;
}
PG3_DEV bool Occluded(const SceneView &S, F3 p, F3 dir, float tmax) {              // scene.hxx:61-74
    const F3 org = p + dir * kEpsRay;
    Hit h;
    h.dist = tmax - 2 * kEpsRay;
    for (uint32_t i = 0; i < S.quadEnd[2]; ++i) {
        const DevQuad q = LoadQuad(S.quads, i);
        if (i < S.quadEnd[0] ? QuadHit<0>(q, org, dir, 0.f, h)
            : i < S.quadEnd[1] ? QuadHit<1>(q, org, dir, 0.f, h) : QuadHit<2>(q, org, dir, 0.f, h))
            return true;
    }
    const F3 m = RayMoment(org, dir);
    for (uint32_t i = 0; i < S.numTris; ++i) {
        const DevTri t = LoadTri(S.tris, i);
        if (TriHit(t, org, dir, m, 0.f, h)) return true;
    }
    for (uint32_t i = 0; i < S.numSphs; ++i) {
        const DevSphere s = LoadSphere(S.sphs, i);
        if (SphHit(s, org, dir, 0.f, h)) return true;
    }
    return false;
}

// Camera::GenerateRay, camera.hxx:91-100 (TransformPoint, math.hxx:734-752)
// One more camera of a scene (multi-view differentiable rendering): raster -> world matrix,
// position, and the scene id keying the RNG streams of its samples.
struct DevView {
    float r2w[16];
    F3 camPos;
    uint32_t sceneId;
};
PG3_DEV void GenerateRay(const float *m, F3 camPos, float x, float y, F3 &org, F3 &dir) {
    float w = m[15];
    w += m[12] * x;
    w += m[13] * y;
    w += m[14] * 0.f;
    const float invW = 1.f / w;
    F3 p;
    p.x = (((m[3] + x * m[0]) + y * m[1]) + 0.f * m[2]) * invW;
    p.y = (((m[7] + x * m[4]) + y * m[5]) + 0.f * m[6]) * invW;
    p.z = (((m[11] + x * m[8]) + y * m[9]) + 0.f * m[10]) * invW;
    org = camPos;
    dir = Normalize(p - camPos);
}
PG3_DEV void GenerateRay(const SceneView &S, float x, float y, F3 &org, F3 &dir) {
    GenerateRay(S.r2w, S.camPos, x, y, org, dir);
}

// ---------------------------------------------------------------------------------------------
// Environment map image (em_image.hxx)
PG3_DEV F3 Texel(const DevEm &E, uint32_t x, uint32_t y) {
    const F4v t = ((GlobalPtr<F4v>)AsGlobal(E.texels))[(size_t)E.w * y + x];
    return mk(t.x, t.y, t.z);
}
PG3_DEV void EmNorm(const DevEm &E, int32_t x, int32_t y, uint32_t &ox, uint32_t &oy) {   // :245-278
    int32_t rx, ry;
    if (y >= (int32_t)E.h) { rx = x + (int32_t)(E.w / 2); ry = (int32_t)E.h - 1; }
    else if (y < 0) { rx = x + (int32_t)(E.w / 2); ry = 0; }
    else { rx = x; ry = y; }
    int32_t m = rx % (int32_t)E.w;
    if (m < 0) m += (int32_t)E.w;
    ox = (uint32_t)m;
    oy = (uint32_t)ry;
}
PG3_DEV F3 EmEvalUV(const DevEm &E, float u, float v) {                             // :95-142
    const float xf = u * (float)E.w, yf = v * (float)E.h;
    if (!E.bilinear) {
        const uint32_t x0 = min((uint32_t)xf, E.w - 1u), y0 = min((uint32_t)yf, E.h - 1u);
        return Texel(E, x0, y0);
    }
    const int32_t cx = (int32_t)(xf + 0.5f), cy = (int32_t)(yf + 0.5f);
    const int32_t x0 = cx - 1, y0 = cy - 1;
    const float xl = xf - ((float)x0 + 0.5f);
    const float yl = yf - ((float)y0 + 0.5f);
    uint32_t ax, ay, bx, by;
    EmNorm(E, x0, y0, ax, ay);
    EmNorm(E, cx, cy, bx, by);
    const F3 a = Texel(E, ax, ay) * (1 - xl) + Texel(E, bx, ay) * xl;
    const F3 b = Texel(E, ax, by) * (1 - xl) + Texel(E, bx, by) * xl;
    return a * (1 - yl) + b * yl;
}
// Texels and weights behind EmEvalUV at a direction, evaluated over any texel array with the
// same arithmetic (differentiable rendering reads the radiance of evalTexels through it and
// scatters texel gradients with the weights).
struct EmFoot {
    uint32_t a, b, c, d;     // texel indices y*w + x of (x0,y0), (x1,y0), (x0,y1), (x1,y1)
    float xl, yl;            // bilinear weights; box filter: a only, xl = yl = 0
};
PG3_DEV EmFoot EmFootprint(const DevEm &E, F3 dir) {
    float u, v;
    Dir2LatLong(dir, u, v);
    const float xf = u * (float)E.w, yf = v * (float)E.h;
    EmFoot f;
    if (!E.bilinear) {
        const uint32_t x0 = min((uint32_t)xf, E.w - 1u), y0 = min((uint32_t)yf, E.h - 1u);
        f.a = f.b = f.c = f.d = E.w * y0 + x0;
        f.xl = f.yl = 0.f;
        return f;
    }
    const int32_t cx = (int32_t)(xf + 0.5f), cy = (int32_t)(yf + 0.5f);
    const int32_t x0 = cx - 1, y0 = cy - 1;
    f.xl = xf - ((float)x0 + 0.5f);
    f.yl = yf - ((float)y0 + 0.5f);
    uint32_t ax, ay, bx, by;
    EmNorm(E, x0, y0, ax, ay);
    EmNorm(E, cx, cy, bx, by);
    f.a = E.w * ay + ax;
    f.b = E.w * ay + bx;
    f.c = E.w * by + ax;
    f.d = E.w * by + bx;
    return f;
}
PG3_DEV F3 TexelAt(const float4 *tex, uint32_t i) {
    const F4v t = ((GlobalPtr<F4v>)AsGlobal(tex))[i];
    return mk(t.x, t.y, t.z);
}
PG3_DEV F3 EmEvalFoot(const float4 *tex, const EmFoot &f) {
    const F3 a = TexelAt(tex, f.a) * (1 - f.xl) + TexelAt(tex, f.b) * f.xl;
    const F3 b = TexelAt(tex, f.c) * (1 - f.xl) + TexelAt(tex, f.d) * f.xl;
    return a * (1 - f.yl) + b * f.yl;
}
// Luminance of EmEvalUV from the luminance map (same texels and weights; luminance is linear).
PG3_DEV float LumAt(const DevEm &E, uint32_t x, uint32_t y) { return AsGlobal(E.lum)[(size_t)E.w * y + x]; }
PG3_DEV float EmLumUV(const DevEm &E, float u, float v) {
    const float xf = u * (float)E.w, yf = v * (float)E.h;
    if (!E.bilinear) {
        const uint32_t x0 = min((uint32_t)xf, E.w - 1u), y0 = min((uint32_t)yf, E.h - 1u);
        return LumAt(E, x0, y0);
    }
    const int32_t cx = (int32_t)(xf + 0.5f), cy = (int32_t)(yf + 0.5f);
    const int32_t x0 = cx - 1, y0 = cy - 1;
    const float xl = xf - ((float)x0 + 0.5f);
    const float yl = yf - ((float)y0 + 0.5f);
    uint32_t ax, ay, bx, by;
    EmNorm(E, x0, y0, ax, ay);
    EmNorm(E, cx, cy, bx, by);
    const float a = LumAt(E, ax, ay) * (1 - xl) + LumAt(E, bx, ay) * xl;
    const float b = LumAt(E, ax, by) * (1 - xl) + LumAt(E, bx, by) * xl;
    return a * (1 - yl) + b * yl;
}
PG3_DEV void Antipode(float u, float v, float &uo, float &vo) {   // lat-long of -dir from that of dir
    uo = (u < 0.5f) ? u + 0.5f : u - 0.5f;
    vo = 1.f - v;
}
PG3_DEV F3 EmEval(const DevEm &E, F3 dir) {
    float u, v;
    Dir2LatLong(dir, u, v);
    return EmEvalUV(E, u, v);
}

// ---------------------------------------------------------------------------------------------
// Simple spherical sampler (em_simple_spherical_sampler.hxx) over the 2-D CDF
// (distribution.hxx:391-455, 510-541): upper_bound over the full CDF selects the same segment as
// the reference's 16-float hierarchical blocks.
template <class P>
PG3_DEV uint32_t UpperBound(P cdf, uint32_t n1, float x) {   // first index with cdf > x
    uint32_t lo = 0, len = n1;
    while (len > 0) {
        const uint32_t half = len >> 1;
        if (cdf[lo + half] <= x) { lo += half + 1; len -= half + 1; }
        else len = half;
    }
    return lo;
}
template <class P>
PG3_DEV void Sample1D(P cdf, uint32_t count, float u, float &x, uint32_t &seg, float &pdf) {
    const float ut = u * 0.999999f;
    int32_t s = (int32_t)UpperBound(cdf, count + 1, ut) - 1;
    s = max(0, min(s, (int32_t)count - 1));
    seg = (uint32_t)s;
    const float c0 = cdf[seg], c1 = cdf[seg + 1];
    const float p = c1 - c0;
    const float off = (ut - c0) / p;
    pdf = p * count;
    x = (seg + off) / count;
}
PG3_DEV float SinMidTheta(uint32_t h, uint32_t sy) { return sinf(kPiF * (sy + 0.5f) / h); }  // :177-189
// margLds (may be null): an LDS copy of the marginal CDF staged by the kernel.
PG3_DEV void SimpleSample(const DevEm &E, const Frame &fr, uint32_t sides, float u0, float u1, F3 &dir,
                          float &pdf, F3 &radCos, const float *margLds = nullptr) {      // :27-81
    float u, v, mp, cp;
    uint32_t sx, sy;
    if (margLds) Sample1D(margLds, E.h, u0, v, sy, mp);
    else Sample1D(AsGlobal(E.margCdf), E.h, u0, v, sy, mp);
    Sample1D(AsGlobal(E.condCdf + (size_t)sy * (E.w + 1)), E.w, u1, u, sx, cp);
    dir = LatLong2Dir(u, v);
    pdf = (mp * cp) * E.plan2ang / SinMidTheta(E.h, sy);
    const F3 rad = !E.bilinear ? Texel(E, sx, sy) : EmEvalUV(E, u, v);
    const float c = Dot(dir, fr.z);
    if (((sides & kSideFront) && c > 0.0f) || ((sides & kSideBack) && c < 0.0f)) radCos = rad * fabsf(c);
    else radCos = mk(0.f);
}
PG3_DEV float SimplePdf(const DevEm &E, F3 dir, const float *margLds = nullptr) {    // :84-98, distribution.hxx:530-540
    float u, v;
    Dir2LatLong(dir, u, v);
    const uint32_t iu = min((uint32_t)fmaxf(0.f, u * E.w), E.w - 1);
    const uint32_t iv = min((uint32_t)fmaxf(0.f, v * E.h), E.h - 1);
    const GlobalPtr<float> cc = AsGlobal(E.condCdf + (size_t)iv * (E.w + 1));
    const float pc = (cc[iu + 1] - cc[iu]) * E.w;
    float m0, m1;
    if (margLds) { m0 = margLds[iv]; m1 = margLds[iv + 1]; }
    else { const GlobalPtr<float> mc = AsGlobal(E.margCdf); m0 = mc[iv]; m1 = mc[iv + 1]; }
    const float pm = (m1 - m0) * E.h;
    const uint32_t seg = min((uint32_t)(v * E.h), E.h - 1u);
    return (pc * pm) * E.plan2ang / SinMidTheta(E.h, seg);
}

// ---------------------------------------------------------------------------------------------
// Steerable sampler run time (em_steerable_sampler.hxx)
struct SH9 {
    float c[9];
};
PG3_DEV SH9 ClampedCos(F3 n) {                                 // GenerateForClampedCos(n, true), :657-688
    const float c1 = .429043f, c2 = .511644f, c3 = .743125f, c4 = .886227f, c5 = .247708f;
    SH9 s;
    s.c[0] = c4 + 0.15f;
    s.c[1] = 2 * c2 * n.y;
    s.c[2] = 2 * c2 * n.z;
    s.c[3] = 2 * c2 * n.x;
    s.c[4] = 2 * c1 * n.x * n.y;
    s.c[5] = 2 * c1 * n.y * n.z;
    s.c[6] = c3 * Sqr(n.z) - c5;
    s.c[7] = 2 * c1 * n.x * n.z;
    s.c[8] = c1 * (Sqr(n.x) - Sqr(n.y));
    return s;
}
template <class P> PG3_DEV float Dot9(P w, const SH9 &c) {            // :135-146 (sequential sum)
    float r = 0.0f;
#pragma unroll
    for (int i = 0; i < 9; ++i) r += w[i] * c.c[i];
    return r;
}
PG3_DEV float ShDotCoeffs(F3 d, const SH9 &c) {               // Dot(GenerateSphHarm(d), coeffs), :181-202
    const float b[9] = {0.282095f, 0.488603f * d.y, 0.488603f * d.z, 0.488603f * d.x,
                        1.092548f * d.x * d.y, 1.092548f * d.y * d.z, 0.315392f * (3.f * d.z * d.z - 1.f),
                        1.092548f * d.x * d.z, 0.546274f * (d.x * d.x - d.y * d.y)};
    float r = 0.0f;
#pragma unroll
    for (int i = 0; i < 9; ++i) r += b[i] * c.c[i];
    return r;
}
PG3_DEV float WholeIntegral(const DevEm &E, const SH9 &c) { return Dot9(E.rootW, c); }  // :2416-2432
// GetBasePdf, :2387-2413, with the EM luminance at dir given
PG3_DEV float BasePdfLum(const DevEm &E, F3 dir, float lum, const SH9 &c) {
    const float cc = ShDotCoeffs(dir, c);
    const float whole = WholeIntegral(E, c);
    return IsTiny(whole) ? 0.f : FastDiv(lum * cc, whole);
}
// EM radiance at a direction and luminance at its antipode with one lat-long mapping: the
// antipode of (u, v) is (u +- 1/2, 1 - v) (the steerable pdf always needs both directions).
struct EmPair {
    F3 at;
    float lumOpp;
};
PG3_DEV EmPair EmEvalPair(const DevEm &E, F3 dir) {
    float u, v, uo, vo;
    Dir2LatLong(dir, u, v);
    Antipode(u, v, uo, vo);
    return EmPair{EmEvalUV(E, u, v), EmLumUV(E, uo, vo)};
}
// Inner levels of the steerable tree staged in LDS by a kernel: levels >= minLevel (down from the
// root) are read from `nodes` (node index - baseNode) instead of E.nodeW.
struct TreeCache {
    const float *nodes = nullptr;
    uint32_t minLevel = 0xFFFFFFFFu;
    uint32_t baseNode = 0;
    const float *marg = nullptr;   // LDS copy of the simple sampler's marginal CDF (or null)
};
// One level of PickTriangle: children [first, first + cc) with weights at w (12 floats per node).
template <class P>
PG3_DEV uint32_t PickChild(P w, uint32_t cc, const SH9 &c, float &u) {
    float cs[7];
    cs[0] = 0.f;
#pragma unroll
    for (uint32_t i = 0; i < 6; ++i) {
        const float ci = (i < cc) ? Dot9(w + i * 12u, c) : 0.f;
        cs[i + 1] = cs[i] + ci;
    }
    float whole = cs[0];
#pragma unroll
    for (uint32_t i = 1; i <= 6; ++i) if (i == cc) whole = cs[i];
    if (whole > 1E-10f) {
        const float inv = 1.f / whole;
#pragma unroll
        for (uint32_t i = 1; i <= 6; ++i) cs[i] *= inv;
    }
    const float ut = u * 0.999999f;
    // upper_bound over cs[0..cc]; past-the-end clamps to the last element
    uint32_t ub = cc;
#pragma unroll
    for (int32_t i = 6; i >= 0; --i) if ((uint32_t)i <= cc && cs[i] > ut) ub = (uint32_t)i;
    if (ub == 0) ub = 1;   // cs[0] = 0 <= ut, unreachable; kept for safety
    float low = 0.f, high = 0.f;
#pragma unroll
    for (uint32_t i = 0; i <= 6; ++i) {
        if (i == ub - 1u) low = cs[i];
        if (i == ub) high = cs[i];
    }
    const float nu = (ut - low) / fmaxf(high - low, 0.00001f);
    u = fminf(nu, 1.f);
    return ub - 1u;
}
// PickTriangle, :3798-3875.  Level l node i has children [6i, min(6i+6, count(l-1))) on level l-1.
PG3_DEV uint32_t PickLeaf(const DevEm &E, const SH9 &c, float &u, const TreeCache &tc = TreeCache{}) {
    uint32_t node = 0;
    for (int32_t lvl = (int32_t)E.numLevels - 1; lvl > 0; --lvl) {
        const uint32_t first = node * 6u;
        const uint32_t cc = min(6u, E.levelCount[lvl - 1] - first);
        const uint32_t idx = E.levelBase[lvl - 1] + first;
        uint32_t child;
        if ((uint32_t)(lvl - 1) >= tc.minLevel)     // wave-uniform: the level is a loop counter
            child = PickChild(tc.nodes + (size_t)(idx - tc.baseNode) * 12u, cc, c, u);
        else
            child = PickChild(AsGlobal(E.nodeW + (size_t)idx * 12u), cc, c, u);
        node = first + child;
    }
    return node;
}
// SampleTriangleFFunction / GFunction / Bilinear, :3709-3747, :3882-3905 (math.hxx:284-342)
PG3_DEV void SampleTriBilinear(float ux, float uy, float va, float vb, float vc, float &b0, float &b1) {
    const float x = (vb - va) / 3.f + (vc - vb) / 6.f;
    const float y = va * 0.5f;
    const float a = x / (x + y), b = y / (x + y);
    float s = 0.5f;
    bool bail = false;
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        const float f = s * (s * (s * a + b) + 0.f) + (-ux);
        const float d = s * (s * 3.f * a + 2.f * b) + 0.f;
        if (!bail) {
            if (fabsf(d) < 0.0001f) { bail = true; }
            else s = Clampf(s - f / d, 0.f, 1.f);
        }
    }
    if (bail) s = 0.5f;
    const float t0 = (s * (vc - vb)) + (2 * (1 - s) * va) + (s * vb);
    const float gamma = (s * (vc - vb)) / t0;
    const float rho = (2 * ((1 - s) * va + s * vb)) / t0;
    const float t = (2 * uy) / (rho + SafeSqrt(rho * rho + 4 * gamma * uy));
    b0 = 1 - s;
    b1 = s * (1 - t);
}
// SampleHemisphereImpl, :2162-2217
// Direction sampled from the linear approximation of a hemisphere (PickTriangle, then a point of
// the leaf triangle with density proportional to the interpolated vertex values) and the
// approximation's value there.
PG3_DEV F3 SteerTriangleSample(const DevEm &E, const SH9 &c, float u0, float u1, const TreeCache &tc, float &val) {
    float ux = u0;
    const uint32_t leaf = PickLeaf(E, c, ux, tc);
    const GlobalPtr<float> lv = AsGlobal(E.leafVerts + (size_t)leaf * 36u);
    const float v0 = Dot9(lv + 3, c), v1 = Dot9(lv + 15, c), v2 = Dot9(lv + 27, c);
    float b0, b1;
    SampleTriBilinear(ux, u1, v0, v1, v2, b0, b1);
    const float b2 = 1.0f - b0 - b1;
    val = b0 * v0 + b1 * v1 + b2 * v2;
    const F3 p = mk(lv[0], lv[1], lv[2]) * b0 + mk(lv[12], lv[13], lv[14]) * b1 + mk(lv[24], lv[25], lv[26]) * b2;
    return FastDiv(p, Length(p));
}
PG3_DEV bool SteerSampleHemisphere(const DevEm &E, F3 n, const SH9 &c, float u0, float u1, F3 &dir,
                                   float &pdf, F3 &radCos, const TreeCache &tc = TreeCache{}) {
    float ux = u0;
    const uint32_t leaf = PickLeaf(E, c, ux, tc);
    const GlobalPtr<float> lv = AsGlobal(E.leafVerts + (size_t)leaf * 36u);
    const float v0 = Dot9(lv + 3, c), v1 = Dot9(lv + 15, c), v2 = Dot9(lv + 27, c);
    float b0, b1;
    SampleTriBilinear(ux, u1, v0, v1, v2, b0, b1);
    const float b2 = 1.0f - b0 - b1;
    const float val = b0 * v0 + b1 * v1 + b2 * v2;
    const F3 p = mk(lv[0], lv[1], lv[2]) * b0 + mk(lv[12], lv[13], lv[14]) * b1 + mk(lv[24], lv[25], lv[26]) * b2;
    dir = FastDiv(p, Length(p));
    const float whole = WholeIntegral(E, c);
    pdf = IsTiny(whole) ? 0.f : FastDiv(val, whole);
    const F3 fl = -dir;
    float cs = Dot(dir, n);
    const bool flip = cs < 0.f;
    if (flip) { dir = fl; cs *= -1.f; }
    // radiance at the returned direction; the base pdf needs the luminance at -(sampled
    // direction), which is the returned direction when it was flipped and its antipode otherwise
    float u, v, uo, vo;
    Dir2LatLong(dir, u, v);
    const F3 Le = EmEvalUV(E, u, v);
    Antipode(u, v, uo, vo);
    pdf += BasePdfLum(E, fl, flip ? Luminance(Le) : EmLumUV(E, uo, vo), c);
    radCos = Le * cs;
    return true;
}
// SampleImpl, :2084-2157.  One call site of the hemisphere sampler: a lane's hemisphere (normal
// and its clamped-cosine coefficients) is selected first, so lanes of a wavefront that pick
// different hemispheres run one tree descent, not two.
PG3_DEV bool SteerSample(const DevEm &E, F3 n, uint32_t sides, float uSide, float u0, float u1, F3 &dir,
                         float &pdf, F3 &radCos, const TreeCache &tc = TreeCache{}) {
    const bool front = sides & kSideFront, back = sides & kSideBack;
    if (!front && !back) return false;
    float ui = 0.f, li = 0.f, wi = 0.f;
    bool up = front;
    if (front && back) {
        ui = WholeIntegral(E, ClampedCos(n));
        li = WholeIntegral(E, ClampedCos(-n));
        wi = ui + li;
        if (IsTiny(wi)) return false;
        up = uSide * wi < ui;
    }
    const F3 hn = up ? n : -n;
    SteerSampleHemisphere(E, hn, ClampedCos(hn), u0, u1, dir, pdf, radCos, tc);
    if (front && back) pdf *= FastDiv(up ? ui : li, wi);
    return true;
}
// pre (may be null): the EM radiance pair of dir, when the caller has it already
PG3_DEV float SteerPdfHemisphere(const DevEm &E, F3 dir, F3 n, const SH9 &c, const EmPair *pre) {  // :2281-2299
    if (Dot(dir, n) < 0.f) return 0.f;
    float la, lo;
    if (pre) {
        la = Luminance(pre->at);
        lo = pre->lumOpp;
    } else {
        float u, v, uo, vo;
        Dir2LatLong(dir, u, v);
        Antipode(u, v, uo, vo);
        la = EmLumUV(E, u, v);
        lo = EmLumUV(E, uo, vo);
    }
    return BasePdfLum(E, dir, la, c) + BasePdfLum(E, -dir, lo, c);
}
PG3_DEV float SteerPdf(const DevEm &E, F3 dir, F3 n, uint32_t sides, const EmPair *pre = nullptr) {  // :2220-2278
    const bool front = sides & kSideFront, back = sides & kSideBack;
    if (!front && !back) return 0.f;
    float ui = 0.f, li = 0.f, wi = 0.f;
    bool up = front;
    if (front && back) {
        ui = WholeIntegral(E, ClampedCos(n));
        li = WholeIntegral(E, ClampedCos(-n));
        wi = ui + li;
        if (IsTiny(wi)) return 0.f;
        up = Dot(dir, n) >= 0.f;
    }
    const F3 hn = up ? n : -n;
    const float p = SteerPdfHemisphere(E, dir, hn, ClampedCos(hn), pre);
    return (front && back) ? p * FastDiv(up ? ui : li, wi) : p;
}
PG3_DEV float SteerIrradiance(const DevEm &E, F3 n, uint32_t sides) {               // :2304-2358
    const float ui = WholeIntegral(E, ClampedCos(n)), li = WholeIntegral(E, ClampedCos(-n));
    const bool front = sides & kSideFront, back = sides & kSideBack;
    if (front && back) return ui + li;
    if (front) return ui;
    if (back) return li;
    return 0.f;
}

// EnvironmentMap::PdfW of the active sampler (em.hxx:113-120)
// `kind` is the sampler (E.sampler at run time, or a constant in specialised kernels).
PG3_DEV float EmPdf(const DevEm &E, int kind, F3 dir, const Frame &fr, uint32_t sides, const EmPair *pre = nullptr,
                     const float *margLds = nullptr) {
    if (kind == kEmSimple) return SimplePdf(E, dir, margLds);
    if (kind == kEmSteerable) return SteerPdf(E, dir, fr.z, sides, pre);
    return CosSpherePdfW(sides & kSideFront, sides & kSideBack, ToLocal(fr, dir));
}
// EnvironmentMapSampler::Sample of the active sampler (em_sampler.hxx:49-64)
PG3_DEV bool EmSample(const DevEm &E, int kind, const Frame &fr, uint32_t sides, float u0, float u1, float u2, F3 &dir,
                      float &pdf, F3 &radCos, const TreeCache &tc = TreeCache{}) {
    if (kind == kEmSimple) { SimpleSample(E, fr, sides, u0, u1, dir, pdf, radCos, tc.marg); return true; }
    if (kind == kEmSteerable) return SteerSample(E, fr.z, sides, u2, u0, u1, dir, pdf, radCos, tc);
    const F3 w = SampleCosSphereParamPdfW(u0, u1, u2, sides & kSideFront, sides & kSideBack, pdf);
    dir = ToWorld(fr, w);
    radCos = EmEval(E, dir) * fabsf(w.z);
    return true;
}

// ---------------------------------------------------------------------------------------------
// Lights (lights.hxx)
struct LightSample {
    F3 sample, wig;
    float pdfW, prob, dist;
};

PG3_DEV float SideCos(uint32_t sides, float c) {                                    // :188-204
    if ((sides & kSideFront) && (sides & kSideBack)) return fabsf(c);
    if (sides & kSideFront) return fmaxf(c, 0.0f);
    if (sides & kSideBack) return fmaxf(-c, 0.0f);
    return 0.0f;
}
// Direction and distance to a light point by the hardware reciprocal square root (v_rsq_f32,
// 1 ulp): light sampling and the light-picking estimates (four of these per area light per
// vertex) need no correctly rounded sqrt and divisions; a one-ulp change moves the shadow ray's
// direction and length by ~1e-7 relative.
PG3_DEV void AreaComputeSample(const DevLight &L, F3 sp, F3 pt, F3 n, uint32_t sides, LightSample &o) {  // :171-219
    o.wig = pt - sp;
    const float d2 = LenSqr(o.wig);
    const float rinv = __builtin_amdgcn_rsqf(d2);
    o.dist = d2 * rinv;
    o.wig = o.wig * rinv;
    const float cosOut = -Dot(L.n, o.wig);
    const float cosIn = SideCos(sides, Dot(n, o.wig));
    o.sample = (cosOut > 0.f) ? L.radiance * cosIn : mk(0.f);
    o.pdfW = fmaxf(L.invArea * FastDiv(d2, fabsf(cosOut)), kEpsDist);
    o.prob = 1.0f;
}
PG3_DEV void PointCompute(const DevLight &L, F3 sp, F3 n, uint32_t sides, LightSample &o) {  // :296-331
    o.wig = L.p0 - sp;
    const float d2 = LenSqr(o.wig);
    o.dist = sqrtf(d2);
    o.wig = o.wig / o.dist;
    const float cosIn = SideCos(sides, Dot(n, o.wig));
    o.sample = (cosIn > 0.f) ? L.radiance * cosIn / d2 : mk(0.f);
    o.pdfW = kInf;
    o.prob = 1.0f;
}
// What a kernel knows about the scene's lights: the EM sampler (kEmAbsent when the background
// has no image) and the light count.  Generic kernels read both from the scene; specialised
// kernels pass constants, which removes the other branches at compile time.
struct LightSpec {
    int em;
    uint32_t numLights;
    TreeCache tree;      // steerable tree levels staged in LDS by the kernel, if any
};
PG3_HD LightSpec SceneLightSpec(const SceneView &S) {
    return LightSpec{S.em.w > 0 ? S.em.sampler : (int)kEmAbsent, S.numLights};
}

// BackgroundLight::GetEmmision(dir), lights.hxx:379-407 (radiance only)
PG3_DEV F3 BgRadiance(const SceneView &S, const LightSpec &ls, F3 dir) {
    return (ls.em != kEmAbsent) ? EmEval(S.em, dir) : LoadGlobal(S.lights + S.bgLight).radiance;
}
// ... and its PDF w.r.t. a shading frame/material
PG3_DEV float BgPdf(const SceneView &S, const LightSpec &ls, F3 dir, const Frame &fr, uint32_t sides,
                    const EmPair *pre = nullptr) {
    if (ls.em != kEmAbsent) return EmPdf(S.em, ls.em, dir, fr, sides, pre, ls.tree.marg);
    return CosSpherePdfW(sides & kSideFront, sides & kSideBack, ToLocal(fr, dir));
}
// AreaLight::GetEmmision PDF part, lights.hxx:88-99
PG3_DEV float AreaEmissionPdf(const DevLight &L, F3 lightPt, F3 surfPt) {
    F3 wig = lightPt - surfPt;
    const float d2 = LenSqr(wig);
    wig = wig * __builtin_amdgcn_rsqf(d2);
    return fmaxf(L.invArea * FastDiv(d2, fabsf(Dot(L.n, wig))), kEpsDist);
}

// EnvironmentMap::EstimateIrradiance MC fallback (em.hxx:123-170), simple + cosine samplers
PG3_DEV float EmEstimateMC(const SceneView &S, const Frame &fr, uint32_t sides, const RngKey &key, uint32_t vtx) {
    const bool front = sides & kSideFront, back = sides & kSideBack;
    float sum = 0.f;
    for (uint32_t round = 0; round < 10; round++) {
        const uint32_t d = kDimEstimateBase + round * 8;
        float p1c;
        const F3 w1 = SampleCosSphereParamPdfW(RngGet(key, vtx, 0, d), RngGet(key, vtx, 0, d + 1),
                                               RngGet(key, vtx, 0, d + 2), front, back, p1c);
        const F3 g1 = ToWorld(fr, w1);
        const F3 s1 = EmEval(S.em, g1) * fabsf(w1.z);
        const float p1s = SimplePdf(S.em, g1);
        F3 g2, s2;
        float p2s;
        SimpleSample(S.em, fr, sides, RngGet(key, vtx, 0, d + 3), RngGet(key, vtx, 0, d + 4), g2, p2s, s2);
        const float p2c = CosSpherePdfW(front, back, ToLocal(fr, g2));
        sum += Luminance(s1) / (p1c + p1s) + Luminance(s2) / (p2c + p2s);
    }
    return sum / 10;
}
// AbstractLight::EstimateContribution
PG3_DEV float EstimateContribution(const SceneView &S, const LightSpec &spec, uint32_t li, F3 sp, const Frame &fr, uint32_t sides,
                                   const RngKey &key, uint32_t vtx) {
    const DevLight L = LoadGlobal(S.lights + li);
    if (L.type == kLightArea) {                                                        // :133-167
        const F3 P1 = L.p0 + L.e1, P2 = L.p0 + L.e2;
        const F3 P3 = L.p0 + 0.33f * L.e1 + 0.33f * L.e2;
        LightSample a, b, c, d;
        AreaComputeSample(L, sp, L.p0, fr.z, sides, a);
        AreaComputeSample(L, sp, P1, fr.z, sides, b);
        AreaComputeSample(L, sp, P2, fr.z, sides, c);
        AreaComputeSample(L, sp, P3, fr.z, sides, d);
        return (FastDiv(Luminance(a.sample), a.pdfW) + FastDiv(Luminance(b.sample), b.pdfW) +
                FastDiv(Luminance(c.sample), c.pdfW) + FastDiv(Luminance(d.sample), d.pdfW)) * 0.25f;
    }
    if (L.type == kLightPoint) {                                                       // :277-290
        LightSample s;
        PointCompute(L, sp, fr.z, sides, s);
        return Luminance(s.sample);
    }
    if (spec.em != kEmAbsent) {                                                        // :451-467
        if (spec.em == kEmSteerable) return SteerIrradiance(S.em, fr.z, sides);
        return EmEstimateMC(S, fr, sides, key, vtx);
    }
    return Luminance(L.radiance) * kPiF;
}
// AbstractLight::SampleIllumination
PG3_DEV bool SampleIllumination(const SceneView &S, const LightSpec &spec, uint32_t li, F3 sp, const Frame &fr, uint32_t sides,
                                float u0, float u1, float u2, LightSample &o) {
    const DevLight L = LoadGlobal(S.lights + li);
    if (L.type == kLightArea) {                                                        // :114-128
        const F3 pt = SampleUniformTriangle(L.p0, L.p0 + L.e1, L.p0 + L.e2, u0, u1);
        AreaComputeSample(L, sp, pt, fr.z, sides, o);
        return true;
    }
    if (L.type == kLightPoint) { PointCompute(L, sp, fr.z, sides, o); return true; }
    o.prob = 1.0f;
    o.dist = 3.402823466e+38f;   // em_sampler.hxx:57
    if (spec.em != kEmAbsent) return EmSample(S.em, spec.em, fr, sides, u0, u1, u2, o.wig, o.pdfW, o.sample, spec.tree);
    const F3 w = SampleCosSphereParamPdfW(u0, u1, u2, sides & kSideFront, sides & kSideBack, o.pdfW);  // em_cosine_sampler.hxx:12-33
    o.wig = ToWorld(fr, w);
    o.sample = L.radiance * fabsf(w.z);
    return true;
}

}  // namespace pg3
