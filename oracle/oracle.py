"""ctypes binding of liboracle.so -- the CPU restatement of the reference hot path.

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, always as the checker, never as the thing measured or shipped.
"""
import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(_HERE))
from pg3render_amd import _ffi  # noqa: E402  (structure definitions shared with the C header)

LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None
_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int32)
_U32P = C.POINTER(C.c_uint32)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"liboracle.so is not built ({LIB_PATH})")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "oracle_scene_create": (C.c_void_p, [C.POINTER(_ffi.SceneDesc)]),
        "oracle_scene_destroy": (None, [C.c_void_p]),
        "oracle_set_params": (None, [C.c_void_p, C.POINTER(_ffi.RenderParams)]),
        "oracle_render": (C.c_double, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_int, C.POINTER(C.c_double),
                                       C.c_uint32, C.c_uint32]),
        "oracle_bsdf_eval": (None, [C.POINTER(_ffi.Material), C.c_uint32, _FP, _FP, _FP, _FP, _FP]),
        "oracle_bsdf_sample": (None, [C.POINTER(_ffi.Material), C.c_uint32, _FP, _FP, _FP, _FP, _FP, _FP]),
        "oracle_em_sample": (None, [C.c_void_p, C.c_uint32, _FP, _IP, _FP, _FP, _FP, _FP, _IP]),
        "oracle_em_pdf": (None, [C.c_void_p, C.c_uint32, _FP, _FP, _IP, _FP]),
        "oracle_em_eval": (None, [C.c_void_p, C.c_uint32, _FP, _FP]),
        "oracle_em_eval_uv": (None, [C.c_void_p, C.c_uint32, _FP, _FP]),
        "oracle_em_avg_pixel_lum": (C.c_float, [C.c_void_p, C.c_uint32, C.c_uint32]),
        "oracle_estimate_contribution": (None, [C.c_void_p, C.c_int32, C.c_uint32, _FP, _FP, _IP, C.c_uint32, _FP]),
        "oracle_camera_ray": (None, [C.c_void_p, C.c_float, C.c_float, _FP, _FP]),
        "oracle_intersect_batch": (None, [C.c_void_p, C.c_uint32, _FP, _FP, _FP, _FP, _FP, _IP]),
        "oracle_occluded_batch": (None, [C.c_void_p, C.c_uint32, _FP, _FP, _FP, _IP]),
        "oracle_sh_basis": (None, [_FP, _FP]),
        "oracle_clamped_cos": (None, [_FP, C.c_int, _FP]),
        "oracle_fresnel_dielectric": (C.c_float, [C.c_float, C.c_float]),
        "oracle_fresnel_conductor": (C.c_float, [C.c_float, C.c_float, C.c_float]),
        "oracle_ggx_d": (C.c_float, [_FP, C.c_float]),
        "oracle_ggx_g1": (C.c_float, [_FP, _FP, C.c_float]),
        "oracle_ggx_vndf_sample": (None, [_FP, C.c_float, C.c_float, C.c_float, _FP]),
        "oracle_ggx_vndf_pdf": (C.c_float, [_FP, _FP, C.c_float]),
        "oracle_halfway_refraction": (None, [_FP, _FP, C.c_float, _FP]),
        "oracle_refract": (None, [_FP, _FP, C.c_float, _FP]),
        "oracle_philox": (None, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, _U32P]),
        "oracle_latlong2dir": (None, [C.c_float, C.c_float, _FP]),
        "oracle_dir2latlong": (None, [_FP, _FP]),
        "oracle_sample_tri_bilinear": (None, [C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, _FP]),
        "oracle_sample_tri_st": (None, [C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, _FP]),
        "oracle_steerable_pdf_parts": (None, [C.c_void_p, _FP, C.c_uint32, _FP, _FP]),
        "oracle_dist1d": (None, [_FP, C.c_uint32, C.c_uint32, _FP, _FP, _U32P, _FP, _FP, _FP]),
        "oracle_steerable_info": (None, [C.c_void_p, _U32P, _U32P]),
        "oracle_steerable_leaves": (None, [C.c_void_p, _FP]),
        "oracle_steerable_root_weight": (None, [C.c_void_p, _FP]),
        "oracle_steerable_pick": (None, [C.c_void_p, _FP, C.c_uint32, _FP, _U32P, _FP]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def fp(a):
    return a.ctypes.data_as(_FP)


def f32(a):
    return np.ascontiguousarray(a, np.float32)


class OracleScene:
    """CPU oracle of one scene description (builds its own EM tables like the reference Init)."""

    def __init__(self, scene_build):
        self.lib = load()
        self.build = scene_build
        self.ptr = self.lib.oracle_scene_create(scene_build.desc_ptr)
        self.res = scene_build.resolution

    def __del__(self):
        try:
            if self.ptr:
                self.lib.oracle_scene_destroy(self.ptr)
        except Exception:
            pass

    def set_params(self, params):
        self.lib.oracle_set_params(self.ptr, C.byref(params))

    def render(self, first, n, threads=0, pixel_begin=0, pixel_end=0):
        """FP64 per-pixel sums of iterations [first, first+n); returns (sums (h,w,3), wall seconds)."""
        w, h = self.res
        out = np.zeros((h, w, 3), np.float64)
        t = self.lib.oracle_render(self.ptr, first, n, threads, out.ctypes.data_as(C.POINTER(C.c_double)),
                                   pixel_begin, pixel_end)
        return out, t

    def em_sample(self, normals, sides, u):
        n = len(normals)
        normals, u = f32(normals), f32(u)
        sides = np.ascontiguousarray(sides, np.int32)
        d = np.zeros((n, 3), np.float32); pdf = np.zeros(n, np.float32)
        rc = np.zeros((n, 3), np.float32); ok = np.zeros(n, np.int32)
        self.lib.oracle_em_sample(self.ptr, n, fp(normals), sides.ctypes.data_as(_IP), fp(u), fp(d), fp(pdf), fp(rc),
                                  ok.ctypes.data_as(_IP))
        return d, pdf, rc, ok

    def em_pdf(self, dirs, normals, sides):
        n = len(dirs)
        dirs, normals = f32(dirs), f32(normals)
        sides = np.ascontiguousarray(sides, np.int32)
        pdf = np.zeros(n, np.float32)
        self.lib.oracle_em_pdf(self.ptr, n, fp(dirs), fp(normals), sides.ctypes.data_as(_IP), fp(pdf))
        return pdf

    def em_eval(self, dirs):
        dirs = f32(dirs)
        rgb = np.zeros((len(dirs), 3), np.float32)
        self.lib.oracle_em_eval(self.ptr, len(dirs), fp(dirs), fp(rgb))
        return rgb

    def em_eval_uv(self, uv):
        uv = f32(uv)
        rgb = np.zeros((len(uv), 3), np.float32)
        self.lib.oracle_em_eval_uv(self.ptr, len(uv), fp(uv), fp(rgb))
        return rgb

    def estimate_contribution(self, light_index, points, normals, sides, iteration=0):
        """Light-picking estimate of one light at surface points (AreaLight::EstimateContribution,
        lights.hxx:133-167; EnvironmentMap::EstimateIrradiance, em.hxx:123-170)."""
        points, normals = f32(points), f32(normals)
        n = len(points)
        sides = np.ascontiguousarray(np.broadcast_to(np.asarray(sides, np.int32), (n,)))
        out = np.zeros(n, np.float32)
        self.lib.oracle_estimate_contribution(self.ptr, light_index, n, fp(points), fp(normals),
                                              sides.ctypes.data_as(_IP), iteration, fp(out))
        return out

    def camera_ray(self, x, y):
        o = np.zeros(3, np.float32); d = np.zeros(3, np.float32)
        self.lib.oracle_camera_ray(self.ptr, x, y, fp(o), fp(d))
        return o, d

    def intersect(self, org, dir, tmin):
        """Scene::Intersect (scene.hxx:42-59): (dist (-1: miss), normal (n, 3), material (-1))."""
        org, dir = f32(org), f32(dir)
        n = len(org)
        tmin = f32(np.broadcast_to(np.asarray(tmin, np.float32), (n,)))
        dist = np.zeros(n, np.float32); nrm = np.zeros((n, 3), np.float32); mat = np.zeros(n, np.int32)
        self.lib.oracle_intersect_batch(self.ptr, n, fp(org), fp(dir), fp(tmin), fp(dist), fp(nrm),
                                        mat.ctypes.data_as(_IP))
        return dist, nrm, mat

    def occluded(self, p, dir, tmax):
        """Scene::Occluded (scene.hxx:61-74): bool per segment."""
        p, dir = f32(p), f32(dir)
        n = len(p)
        tmax = f32(np.broadcast_to(np.asarray(tmax, np.float32), (n,)))
        occ = np.zeros(n, np.int32)
        self.lib.oracle_occluded_batch(self.ptr, n, fp(p), fp(dir), fp(tmax), occ.ctypes.data_as(_IP))
        return occ != 0

    def steerable(self):
        nl, nv = C.c_uint32(), C.c_uint32()
        self.lib.oracle_steerable_info(self.ptr, C.byref(nl), C.byref(nv))
        leaves = np.zeros((nl.value, 3, 3), np.float32)
        self.lib.oracle_steerable_leaves(self.ptr, fp(leaves))
        w = np.zeros(9, np.float32)
        self.lib.oracle_steerable_root_weight(self.ptr, fp(w))
        return leaves, int(nv.value), w

    def steerable_pdf_parts(self, normal, dirs):
        """(GetBasePdf(d), GetBasePdf(-d), ComputeWholeIntegral) per direction for a normal's lobe."""
        dirs = f32(dirs)
        out = np.zeros((len(dirs), 3), np.float32)
        self.lib.oracle_steerable_pdf_parts(self.ptr, fp(f32(normal)), len(dirs), fp(dirs), fp(out))
        return out

    def steerable_pick(self, normal, u):
        u = f32(u)
        leaf = np.zeros(len(u), np.uint32)
        nl, nv = C.c_uint32(), C.c_uint32()
        self.lib.oracle_steerable_info(self.ptr, C.byref(nl), C.byref(nv))
        rel = np.zeros(nl.value, np.float32)
        self.lib.oracle_steerable_pick(self.ptr, fp(f32(normal)), len(u), fp(u), leaf.ctypes.data_as(_U32P), fp(rel))
        return leaf, rel


def bsdf_eval(mat, wil, wol):
    lib = load()
    n = len(wil)
    wil, wol = f32(wil), f32(wol)
    f = np.zeros((n, 3), np.float32); pdf = np.zeros(n, np.float32); comp = np.zeros(n, np.float32)
    lib.oracle_bsdf_eval(C.byref(mat), n, fp(wil), fp(wol), fp(f), fp(pdf), fp(comp))
    return f, pdf, comp


def bsdf_sample(mat, wol, u):
    lib = load()
    n = len(wol)
    wol, u = f32(wol), f32(u)
    wil = np.zeros((n, 3), np.float32); f = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32); comp = np.zeros(n, np.float32)
    lib.oracle_bsdf_sample(C.byref(mat), n, fp(wol), fp(u), fp(wil), fp(f), fp(pdf), fp(comp))
    return wil, f, pdf, comp


def v3(x):
    return f32(x)


def sh_basis(d):
    out = np.zeros(9, np.float32); load().oracle_sh_basis(fp(v3(d)), fp(out)); return out


def clamped_cos(n, comp):
    out = np.zeros(9, np.float32); load().oracle_clamped_cos(fp(v3(n)), int(comp), fp(out)); return out


def fresnel_dielectric(c, eta):
    return load().oracle_fresnel_dielectric(c, eta)


def fresnel_conductor(c, eta, k):
    return load().oracle_fresnel_conductor(c, eta, k)


def ggx_d(h, a):
    return load().oracle_ggx_d(fp(v3(h)), a)


def ggx_g1(v, h, a):
    return load().oracle_ggx_g1(fp(v3(v)), fp(v3(h)), a)


def vndf_sample(wo, a, u0, u1):
    h = np.zeros(3, np.float32); load().oracle_ggx_vndf_sample(fp(v3(wo)), a, u0, u1, fp(h)); return h


def vndf_pdf(wo, h, a):
    return load().oracle_ggx_vndf_pdf(fp(v3(wo)), fp(v3(h)), a)


def halfway_refraction(wi, wo, eta):
    h = np.zeros(3, np.float32); load().oracle_halfway_refraction(fp(v3(wi)), fp(v3(wo)), eta, fp(h)); return h


def refract(d, n, eta):
    o = np.zeros(3, np.float32); load().oracle_refract(fp(v3(d)), fp(v3(n)), eta, fp(o)); return o


def philox(c0, c1, c2, c3, seed):
    out = np.zeros(4, np.uint32)
    load().oracle_philox(c0, c1, c2, c3, seed, out.ctypes.data_as(_U32P))
    return out


def latlong2dir(u, v):
    d = np.zeros(3, np.float32); load().oracle_latlong2dir(u, v, fp(d)); return d


def dir2latlong(d):
    uv = np.zeros(2, np.float32); load().oracle_dir2latlong(fp(v3(d)), fp(uv)); return uv


def sample_tri_bilinear(ux, uy, a, b, c):
    o = np.zeros(2, np.float32); load().oracle_sample_tri_bilinear(ux, uy, a, b, c, fp(o)); return o


def sample_tri_st(ux, uy, a, b, c):
    o = np.zeros(2, np.float32); load().oracle_sample_tri_st(ux, uy, a, b, c, fp(o)); return o


def dist1d(f, u):
    f = f32(f); u = f32(u)
    n, m = len(f), len(u)
    x = np.zeros(m, np.float32); seg = np.zeros(m, np.uint32); pdf = np.zeros(m, np.float32)
    cdf = np.zeros(n + 1, np.float32); cdf32 = np.zeros(n + 1, np.float32)
    load().oracle_dist1d(fp(f), n, m, fp(u), fp(x), seg.ctypes.data_as(_U32P), fp(pdf), fp(cdf), fp(cdf32))
    return x, seg, pdf, cdf, cdf32
