"""Host-side mirror of the reference renderer interface for the GPU path.

`GpuPathTracer` has the shape of `AbstractRenderer` (src/renderer.hxx:232-278):
`run_iteration()` is `RunIteration` (one sample per pixel, src/path_tracer_base.hxx:57-90),
`get_framebuffer()` is `GetFramebuffer` (FP64 sum / iterations, src/renderer.hxx:255-261) and
`was_used()` is `WasUsed`.  Scenes come from the C++ builders in libpg3g (include/pg3g_scenes.h).
"""
import ctypes as C
import os
import weakref

import numpy as np

from . import _ffi
from ._ffi import check


class SceneBuild:
    """Owns a host scene description built by libpg3g (workload or reference scene ID)."""

    def __init__(self, handle):
        self._lib = _ffi.load()
        self.handle = handle

    @classmethod
    def workload(cls, workload, seed=1, variant=0, res=0):
        lib = _ffi.load()
        h = C.c_void_p()
        check(lib.pg3g_build_workload_scene(workload, seed, variant, res, C.byref(h)), "pg3g_build_workload_scene")
        return cls(h)

    @classmethod
    def reference(cls, scene_id, em_type=0, resolution=0, em_path=None):
        """Reference scene ID; em_path names the image of an EM type 2..12 (.hdr, PFM or raw)."""
        lib = _ffi.load()
        h = C.c_void_p()
        if em_path is None:
            check(lib.pg3g_build_reference_scene(scene_id, em_type, resolution, C.byref(h)), "pg3g_build_reference_scene")
        else:
            check(lib.pg3g_build_reference_scene_em(scene_id, em_type, resolution, os.fsencode(em_path), C.byref(h)),
                  "pg3g_build_reference_scene_em")
        return cls(h)

    @property
    def desc(self):
        return self._lib.pg3g_scene_build_mutable(self.handle).contents

    @property
    def desc_ptr(self):
        return self._lib.pg3g_scene_build_mutable(self.handle)

    @property
    def spp(self):
        return int(self._lib.pg3g_scene_build_spp(self.handle))

    @property
    def resolution(self):
        c = self.desc.camera
        return int(c.res_x), int(c.res_y)

    def em_texels(self):
        """numpy view (h, w, 3) of the EM image (mutable), or None."""
        d = self.desc
        if d.em.width == 0:
            return None
        p = self._lib.pg3g_scene_build_em_texels(self.handle)
        return np.ctypeslib.as_array(p, shape=(d.em.height, d.em.width, 3))

    def materials(self):
        d = self.desc
        return [d.materials[i] for i in range(d.num_materials)]

    def set_params(self, **kw):
        p = self.desc.params
        for k, v in kw.items():
            if k == "algorithm" and isinstance(v, str):
                v = _ffi.ALGORITHM_ACRONYMS[v]
            setattr(p, k, v)

    def close(self):
        if self.handle:
            self._lib.pg3g_scene_build_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def steerable_cache_dir(path):
    """Sets the directory of the on-disk steerable-table cache (None: no cache)."""
    check(_ffi.load().pg3g_steerable_cache_dir(None if path is None else os.fsencode(path)), "pg3g_steerable_cache_dir")


def load_image(path, em_rotation=None, em_scale=1.0):
    """RGB float32 image (h, w, 3) from a Radiance .hdr, PFM or raw float file; with em_rotation
    set, rotated in azimuth and scaled as the reference loads an environment map."""
    lib = _ffi.load()
    w, h = C.c_uint32(), C.c_uint32()
    p = os.fsencode(path)
    if em_rotation is None:
        check(lib.pg3g_load_image(p, C.byref(w), C.byref(h), None, 0), "pg3g_load_image")
        img = np.zeros((h.value, w.value, 3), np.float32)
        check(lib.pg3g_load_image(p, C.byref(w), C.byref(h), img.ctypes.data_as(C.POINTER(C.c_float)), img.size),
              "pg3g_load_image")
    else:
        check(lib.pg3g_load_em(p, em_rotation, em_scale, C.byref(w), C.byref(h), None, 0), "pg3g_load_em")
        img = np.zeros((h.value, w.value, 3), np.float32)
        check(lib.pg3g_load_em(p, em_rotation, em_scale, C.byref(w), C.byref(h),
                               img.ctypes.data_as(C.POINTER(C.c_float)), img.size), "pg3g_load_em")
    return img


def save_image(path, rgb, scale=1.0, gamma=1.0):
    """Writes an (h, w, 3) image as .hdr (RGBE) or .bmp (24-bit, value^(1/gamma))."""
    a = np.ascontiguousarray(rgb, np.float64)
    check(_ffi.load().pg3g_save_image(os.fsencode(path), a.ctypes.data_as(C.POINTER(C.c_double)), a.shape[1],
                                      a.shape[0], scale, gamma), "pg3g_save_image")


def compare_images(image, reference):
    """Per-pixel and whole-image relative RMSE of luminance (pg3g_image_compare) and |diff|."""
    a = np.ascontiguousarray(image, np.float32)
    b = np.ascontiguousarray(reference, np.float32)
    if a.shape != b.shape:
        raise ValueError(f"image shapes differ: {a.shape} vs {b.shape}")
    d = _ffi.ImageDiff()
    diff = np.zeros_like(a)
    fp = C.POINTER(C.c_float)
    check(_ffi.load().pg3g_image_compare(a.ctypes.data_as(fp), b.ctypes.data_as(fp), a.shape[1], a.shape[0],
                                         C.byref(d), diff.ctypes.data_as(fp)), "pg3g_image_compare")
    return {k: float(getattr(d, k)) for k, _ in _ffi.ImageDiff._fields_}, diff


def steerable_build(em_desc, threads=0, leaves=True):
    """Host-only build of the steerable sampler tables of `em_desc` (an EmDesc, e.g.
    SceneBuild.desc.em): returns (info dict, leaf directions (n, 3, 3) or None, root weight (9,))."""
    lib = _ffi.load()
    info = _ffi.SteerableBuildInfo()
    root = np.zeros(9, np.float32)
    check(lib.pg3g_steerable_build(C.byref(em_desc), threads, C.byref(info), None, 0,
                                   root.ctypes.data_as(C.POINTER(C.c_float))), "pg3g_steerable_build")
    from_cache, build_ms = bool(info.from_cache), float(info.build_ms)   # of the first call
    out = None
    if leaves:
        out = np.zeros((info.num_leaves, 3, 3), np.float32)
        check(lib.pg3g_steerable_build(C.byref(em_desc), threads, C.byref(info),
                                       out.ctypes.data_as(C.POINTER(C.c_float)), info.num_leaves, None),
              "pg3g_steerable_build")
    d = {"num_leaves": int(info.num_leaves), "num_vertices": int(info.num_vertices),
         "num_levels": int(info.num_levels), "threads": int(info.threads), "from_cache": from_cache,
         "build_ms": build_ms}
    return d, out, root


class Context:
    """A libpg3g context on one HIP device."""

    def __init__(self, device=0):
        self._lib = _ffi.load()
        self.ptr = C.c_void_p()
        self._batches = weakref.WeakSet()   # closed before the context (they refer to it)
        check(self._lib.pg3g_create(C.byref(self.ptr), device), "pg3g_create")

    def close(self):
        if self.ptr:
            for b in list(self._batches):
                b.close()
            self._lib.pg3g_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, scene, slot=0):
        check(self._lib.pg3g_upload_scene_slot(self.ptr, slot, scene.desc_ptr), "pg3g_upload_scene_slot")

    def set_params(self, slot, params):
        check(self._lib.pg3g_set_params(self.ptr, slot, C.byref(params)), "pg3g_set_params")

    def set_kernel_mode(self, mode):
        check(self._lib.pg3g_set_kernel_mode(self.ptr, mode), "pg3g_set_kernel_mode")

    def render(self, first_iter, n_iter, slot=0):
        check(self._lib.pg3g_render_slot(self.ptr, slot, first_iter, n_iter), "pg3g_render_slot")

    def run_iteration(self, iteration, slot=0, launch_now=False):
        """Queues one RunIteration (pg3g_run_iteration): consecutive iterations launch together."""
        check(self._lib.pg3g_run_iteration(self.ptr, slot, iteration, 1 if launch_now else 0), "pg3g_run_iteration")

    def render_timed(self, first_iter, n_iter, slot=0):
        """Renders and returns the GPU time (ms) between HIP events on the library's stream."""
        ms = C.c_float()
        check(self._lib.pg3g_render_timed(self.ptr, slot, first_iter, n_iter, C.byref(ms)), "pg3g_render_timed")
        return float(ms.value)

    def synchronize(self):
        check(self._lib.pg3g_synchronize(self.ptr), "pg3g_synchronize")

    def framebuffer(self, slot, res):
        w, h = res
        out = np.zeros((h, w, 3), dtype=np.float64)
        it = C.c_uint64()
        check(self._lib.pg3g_read_framebuffer_slot(self.ptr, slot, out.ctypes.data_as(C.POINTER(C.c_double)),
                                                   C.byref(it)), "pg3g_read_framebuffer_slot")
        return out, int(it.value)

    def clear(self, slot=0):
        check(self._lib.pg3g_clear_framebuffer(self.ptr, slot), "pg3g_clear_framebuffer")

    def counters(self):
        s, c = C.c_uint64(), C.c_uint64()
        check(self._lib.pg3g_read_counters(self.ptr, C.byref(s), C.byref(c)), "pg3g_read_counters")
        return int(s.value), int(c.value)

    def ray_counters(self):
        """(extension rays, shadow rays) traced since the last reset_counters()."""
        e, sh = C.c_uint64(), C.c_uint64()
        check(self._lib.pg3g_read_ray_counters(self.ptr, C.byref(e), C.byref(sh)), "pg3g_read_ray_counters")
        return int(e.value), int(sh.value)

    def primitives(self, slot=0):
        """(rectangles, triangles, spheres) of a slot as the intersection kernels test them."""
        r, t, s = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(self._lib.pg3g_scene_primitives(self.ptr, slot, C.byref(r), C.byref(t), C.byref(s)),
              "pg3g_scene_primitives")
        return int(r.value), int(t.value), int(s.value)

    def reset_counters(self):
        check(self._lib.pg3g_reset_counters(self.ptr), "pg3g_reset_counters")

    # ---- function-level entry points --------------------------------------------------------
    def bsdf_eval(self, mat, wil, wol):
        n = len(wil)
        wil = np.ascontiguousarray(wil, np.float32)
        wol = np.ascontiguousarray(wol, np.float32)
        f = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32)
        comp = np.zeros(n, np.float32)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        check(self._lib.pg3g_bsdf_eval_batch(self.ptr, C.byref(mat), n, fp(wil), fp(wol), fp(f), fp(pdf), fp(comp)),
              "pg3g_bsdf_eval_batch")
        return f, pdf, comp

    def bsdf_sample(self, mat, wol, u):
        n = len(wol)
        wol = np.ascontiguousarray(wol, np.float32)
        u = np.ascontiguousarray(u, np.float32)
        wil = np.zeros((n, 3), np.float32)
        f = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32)
        comp = np.zeros(n, np.float32)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        check(self._lib.pg3g_bsdf_sample_batch(self.ptr, C.byref(mat), n, fp(wol), fp(u), fp(wil), fp(f), fp(pdf),
                                               fp(comp)), "pg3g_bsdf_sample_batch")
        return wil, f, pdf, comp

    def em_sample(self, normals, sides, u):
        n = len(normals)
        normals = np.ascontiguousarray(normals, np.float32)
        sides = np.ascontiguousarray(sides, np.int32)
        u = np.ascontiguousarray(u, np.float32)
        d = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32)
        rc = np.zeros((n, 3), np.float32)
        ok = np.zeros(n, np.int32)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        check(self._lib.pg3g_em_sample_batch(self.ptr, n, fp(normals), ip(sides), fp(u), fp(d), fp(pdf), fp(rc), ip(ok)),
              "pg3g_em_sample_batch")
        return d, pdf, rc, ok

    def em_pdf(self, dirs, normals, sides):
        n = len(dirs)
        dirs = np.ascontiguousarray(dirs, np.float32)
        normals = np.ascontiguousarray(normals, np.float32)
        sides = np.ascontiguousarray(sides, np.int32)
        pdf = np.zeros(n, np.float32)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        check(self._lib.pg3g_em_pdf_batch(self.ptr, n, fp(dirs), fp(normals),
                                          sides.ctypes.data_as(C.POINTER(C.c_int32)), fp(pdf)), "pg3g_em_pdf_batch")
        return pdf

    def em_eval(self, dirs):
        n = len(dirs)
        dirs = np.ascontiguousarray(dirs, np.float32)
        rgb = np.zeros((n, 3), np.float32)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        check(self._lib.pg3g_em_eval_batch(self.ptr, n, fp(dirs), fp(rgb)), "pg3g_em_eval_batch")
        return rgb

    def intersect(self, org, dir, tmin):
        """Scene::Intersect of slot 0 for n rays: (dist (-1: miss), normal (n, 3), material (-1))."""
        n = len(org)
        org = np.ascontiguousarray(org, np.float32)
        dir = np.ascontiguousarray(dir, np.float32)
        tmin = np.ascontiguousarray(np.broadcast_to(np.asarray(tmin, np.float32), (n,)))
        dist = np.zeros(n, np.float32)
        nrm = np.zeros((n, 3), np.float32)
        mat = np.zeros(n, np.int32)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        check(self._lib.pg3g_intersect_batch(self.ptr, n, fp(org), fp(dir), fp(tmin), fp(dist), fp(nrm),
                                             mat.ctypes.data_as(C.POINTER(C.c_int32))), "pg3g_intersect_batch")
        return dist, nrm, mat

    def occluded(self, p, dir, tmax):
        """Scene::Occluded of slot 0 for n segments: bool array."""
        n = len(p)
        p = np.ascontiguousarray(p, np.float32)
        dir = np.ascontiguousarray(dir, np.float32)
        tmax = np.ascontiguousarray(np.broadcast_to(np.asarray(tmax, np.float32), (n,)))
        occ = np.zeros(n, np.int32)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        check(self._lib.pg3g_occluded_batch(self.ptr, n, fp(p), fp(dir), fp(tmax),
                                            occ.ctypes.data_as(C.POINTER(C.c_int32))), "pg3g_occluded_batch")
        return occ != 0

    def steerable_leaves(self):
        nl, nv, lv = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(self._lib.pg3g_steerable_info(self.ptr, C.byref(nl), C.byref(nv), C.byref(lv)), "pg3g_steerable_info")
        out = np.zeros((nl.value, 3, 3), np.float32)
        check(self._lib.pg3g_steerable_leaves(self.ptr, out.ctypes.data_as(C.POINTER(C.c_float))), "pg3g_steerable_leaves")
        w = np.zeros(9, np.float32)
        check(self._lib.pg3g_steerable_root_weight(self.ptr, w.ctypes.data_as(C.POINTER(C.c_float))),
              "pg3g_steerable_root_weight")
        return out, int(nv.value), int(lv.value), w


class GpuPathTracer:
    """AbstractRenderer-shaped GPU renderer for one scene (src/renderer.hxx:232-278)."""

    def __init__(self, scene, context=None, slot=0):
        self.scene = scene
        self.ctx = context or Context()
        self.slot = slot
        self.ctx.upload(scene, slot)
        self.iterations = 0

    def run_iteration(self, iteration, last=False):
        """One sample per pixel (PathTracerBase::RunIteration, src/path_tracer_base.hxx:57-90),
        queued: consecutive iterations launch together (pg3g_run_iteration); `last` launches what is
        queued at once.  get_framebuffer launches whatever is still queued."""
        self.ctx.run_iteration(iteration, self.slot, launch_now=last)
        self.iterations += 1

    def run_iterations(self, first, count):
        self.ctx.render(first, count, self.slot)
        self.iterations += count

    def get_framebuffer(self):
        """FP64 image = sum / iterations (src/renderer.hxx:255-261), shape (h, w, 3)."""
        fb, it = self.ctx.framebuffer(self.slot, self.scene.resolution)
        return fb / it if it > 0 else fb

    def was_used(self):
        return self.iterations > 0


class Batch:
    """Many uploaded slots rendered by one persistent launch per pass (pg3g_batch_*).

    Pass p renders iterations [p*spp[i], (p+1)*spp[i]) of scene i into its slot's framebuffer,
    sample for sample what `Context.render(p*spp[i], spp[i], slot)` renders.
    """

    def __init__(self, ctx, slots, spp):
        self._lib = _ffi.load()
        self.ctx = ctx
        self.slots = [int(s) for s in slots]
        self.spp = [int(n) for n in spp]
        n = len(self.slots)
        sl = (C.c_uint32 * n)(*self.slots)
        sp = (C.c_uint32 * n)(*self.spp)
        self.ptr = C.c_void_p()
        check(self._lib.pg3g_batch_create(ctx.ptr, n, sl, sp, C.byref(self.ptr)), "pg3g_batch_create")
        ctx._batches.add(self)

    def close(self):
        if self.ptr:
            self._lib.pg3g_batch_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def config(self):
        c = _ffi.BatchConfig()
        check(self._lib.pg3g_batch_get_config(self.ptr, C.byref(c)), "pg3g_batch_get_config")
        return {"iters_per_item": c.iters_per_item, "waves_per_cu": c.waves_per_cu, "use_graph": c.use_graph}

    def set_config(self, iters_per_item=8, waves_per_cu=16, use_graph=True):
        c = _ffi.BatchConfig(iters_per_item, waves_per_cu, 1 if use_graph else 0)
        check(self._lib.pg3g_batch_set_config(self.ptr, C.byref(c)), "pg3g_batch_set_config")

    def render(self, passes=1):
        check(self._lib.pg3g_batch_render(self.ptr, passes), "pg3g_batch_render")

    def render_timed(self, passes=1):
        ms = C.c_float()
        check(self._lib.pg3g_batch_render_timed(self.ptr, passes, C.byref(ms)), "pg3g_batch_render_timed")
        return float(ms.value)

    def autotune(self, max_candidates=0):
        ms = C.c_float()
        check(self._lib.pg3g_batch_autotune(self.ptr, max_candidates, C.byref(ms)), "pg3g_batch_autotune")
        return float(ms.value)
