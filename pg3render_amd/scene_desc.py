"""Building a pg3g_scene_desc from Python arrays (include/pg3g.h).

`CustomScene` keeps the numpy/ctypes buffers alive for as long as the description is used, and
exposes the same interface as `SceneBuild` (desc, desc_ptr, resolution, spp, em_texels).
"""
import ctypes as C

import numpy as np

from . import _ffi


def material(kind, diffuse=(0, 0, 0), glossy=(1, 1, 1), alpha=0.1, inner_ior=1.0, outer_ior=1.0, absorbance=0.0,
             phong_exponent=1.0, phong_diffuse_on=1, phong_glossy_on=0, light_id=-1, back_side_light_sampling=0):
    m = _ffi.Material()
    m.type = kind
    m.light_id = light_id
    m.diffuse = _ffi.F3(*diffuse)
    m.glossy = _ffi.F3(*glossy)
    m.phong_exponent = phong_exponent
    m.phong_diffuse_on = phong_diffuse_on
    m.phong_glossy_on = phong_glossy_on
    m.alpha = alpha
    m.inner_ior = inner_ior
    m.outer_ior = outer_ior
    m.absorbance = absorbance
    m.back_side_light_sampling = back_side_light_sampling
    return m


def light(kind, p0=(0, 0, 0), p1=(0, 0, 0), p2=(0, 0, 0), power=(0, 0, 0)):
    lt = _ffi.Light()
    lt.type = kind
    lt.p0 = _ffi.F3(*p0)
    lt.p1 = _ffi.F3(*p1)
    lt.p2 = _ffi.F3(*p2)
    lt.power = _ffi.F3(*power)
    return lt


class CustomScene:
    def __init__(self, triangles=(), triangle_mat=(), spheres=(), sphere_mat=(), materials=(), lights=(),
                 camera=None, em=None, em_sampler=_ffi.EM_SIMPLE, em_bilinear=1, steerable_levels=(4, 4),
                 steerable_max_error=5.0, params=None, spp=1):
        self._tris = np.ascontiguousarray(np.asarray(triangles, np.float32).reshape(-1, 9))
        self._tmat = np.ascontiguousarray(np.asarray(triangle_mat, np.int32).reshape(-1))
        self._sph = np.ascontiguousarray(np.asarray(spheres, np.float32).reshape(-1, 4))
        self._smat = np.ascontiguousarray(np.asarray(sphere_mat, np.int32).reshape(-1))
        self._mats = (_ffi.Material * max(1, len(materials)))(*materials)
        self._lights = (_ffi.Light * max(1, len(lights)))(*lights)
        self._em = None if em is None else np.ascontiguousarray(np.asarray(em, np.float32))
        d = _ffi.SceneDesc()
        d.num_triangles = len(self._tmat)
        d.triangles = self._tris.ctypes.data_as(C.POINTER(C.c_float))
        d.triangle_mat = self._tmat.ctypes.data_as(C.POINTER(C.c_int32))
        d.num_spheres = len(self._smat)
        d.spheres = self._sph.ctypes.data_as(C.POINTER(C.c_float))
        d.sphere_mat = self._smat.ctypes.data_as(C.POINTER(C.c_int32))
        d.num_materials = len(materials)
        d.materials = C.cast(self._mats, C.POINTER(_ffi.Material))
        d.num_lights = len(lights)
        d.lights = C.cast(self._lights, C.POINTER(_ffi.Light))
        if self._em is not None:
            h, w, _ = self._em.shape
            d.em.width = w
            d.em.height = h
            d.em.texels = self._em.ctypes.data_as(C.POINTER(C.c_float))
            d.em.bilinear = em_bilinear
            d.em.sampler = em_sampler
            d.em.st_max_approx_error = steerable_max_error
            d.em.st_min_subdiv_level, d.em.st_max_subdiv_level = steerable_levels
        cam = camera or {}
        d.camera.position = _ffi.F3(*cam.get("position", (0.5, -1.25, 0.5)))
        d.camera.forward = _ffi.F3(*cam.get("forward", (0.0, 1.0, 0.0)))
        d.camera.up = _ffi.F3(*cam.get("up", (0.0, 0.0, 1.0)))
        d.camera.hfov_deg = cam.get("hfov", 45.0)
        d.camera.res_x, d.camera.res_y = cam.get("res", (64, 64))
        p = params or {}
        d.params.algorithm = p.get("algorithm", _ffi.ALG_PATH_TRACING)
        d.params.min_path_length = p.get("min_path_length", 1)
        d.params.max_path_length = p.get("max_path_length", 8)
        d.params.indirect_clipping = p.get("indirect_clipping", 0.0)
        d.params.splitting_budget = p.get("splitting_budget", 1.0)
        d.params.splitting_level = p.get("splitting_level", 1.0)
        d.params.light_to_bsdf_ratio = p.get("light_to_bsdf_ratio", 1.0)
        d.params.seed = p.get("seed", 1234)
        d.params.scene_id = p.get("scene_id", 0)
        self._desc = d
        self.spp = spp

    @property
    def desc(self):
        return self._desc

    @property
    def desc_ptr(self):
        return C.pointer(self._desc)

    @property
    def resolution(self):
        return int(self._desc.camera.res_x), int(self._desc.camera.res_y)

    def em_texels(self):
        return self._em

    def set_params(self, **kw):
        for k, v in kw.items():
            if k == "algorithm" and isinstance(v, str):
                v = _ffi.ALGORITHM_ACRONYMS[v]
            setattr(self._desc.params, k, v)
