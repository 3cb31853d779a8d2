"""Differentiable rendering through libpg3g (pg3g_diff_*, pg3g_render_diff, pg3g_adjoint).

`DiffScene` drives one uploaded slot: the differentiated parameters are the albedo of Lambert
materials (RGB), the tint (RGB) and roughness alpha of GGX conductors, and the EM texels.  The
forward render uses frozen sampling (directions and pdfs from the sampling parameters, values
from the evaluation parameters) and the adjoint is path-replay backpropagation on the GPU with
FP64 accumulation of the EM-texel gradients.

`render_function` wraps a scene as a `torch.autograd.Function`: forward renders the image for the
given parameters (which also become the sampling parameters), backward runs the adjoint of the
same samples.
"""
import ctypes as C

import numpy as np

from . import _ffi
from ._ffi import check


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class DiffScene:
    """Differentiable view of slot `slot` of a Context (the scene must already be uploaded)."""

    def __init__(self, ctx, scene, material_ids, slot=0):
        self._lib = _ffi.load()
        self.ctx, self.scene, self.slot = ctx, scene, slot
        ids = (C.c_uint32 * len(material_ids))(*[int(m) for m in material_ids])
        n = C.c_uint32()
        check(self._lib.pg3g_diff_setup(ctx.ptr, slot, len(material_ids), ids, C.byref(n)), "pg3g_diff_setup")
        self.num_scalars = int(n.value)
        d = scene.desc
        self.em_shape = (int(d.em.height), int(d.em.width), 3) if d.em.width else None
        self.resolution = scene.resolution

    def get_params(self):
        sc = np.zeros(max(self.num_scalars, 1), np.float32)
        em = np.zeros(self.em_shape, np.float32) if self.em_shape else None
        check(self._lib.pg3g_diff_get_params(self.ctx.ptr, self.slot, _fp(sc), _fp(em) if em is not None else None),
              "pg3g_diff_get_params")
        return sc[:self.num_scalars], em

    def set_params(self, scalars=None, em=None, refresh_sampling=True):
        sc = None if scalars is None else np.ascontiguousarray(scalars, np.float32)
        if sc is not None and sc.size != self.num_scalars:
            raise ValueError(f"expected {self.num_scalars} scalars, got {sc.size}")
        e = None if em is None else np.ascontiguousarray(em, np.float32)
        if e is not None and e.shape != self.em_shape:
            raise ValueError(f"expected EM of shape {self.em_shape}, got {e.shape}")
        check(self._lib.pg3g_diff_set_params(self.ctx.ptr, self.slot, _fp(sc) if sc is not None else None,
                                             _fp(e) if e is not None else None, 1 if refresh_sampling else 0),
              "pg3g_diff_set_params")

    def render(self, first_iter, n_iter, timed=False):
        """Mean image (H, W, 3) FP64 of iterations [first_iter, first_iter + n_iter)."""
        self.ctx.clear(self.slot)
        ms = C.c_float()
        check(self._lib.pg3g_render_diff(self.ctx.ptr, self.slot, first_iter, n_iter, C.byref(ms) if timed else None),
              "pg3g_render_diff")
        img, it = self.ctx.framebuffer(self.slot, self.resolution)
        img = img / max(it, 1)
        return (img, float(ms.value)) if timed else img

    def adjoint(self, first_iter, n_iter, dloss_dimage, timed=False):
        """Gradients (scalars FP64, EM texels FP64 (H, W, 3) or None) of a loss whose derivative
        with respect to the mean image of the same iterations is `dloss_dimage` (H, W, 3)."""
        w, h = self.resolution
        adj = np.ascontiguousarray(dloss_dimage, np.float32).reshape(h, w, 3)
        gs = np.zeros(max(self.num_scalars, 1), np.float64)
        ge = np.zeros(self.em_shape, np.float64) if self.em_shape else None
        ms = C.c_float()
        check(self._lib.pg3g_adjoint(self.ctx.ptr, self.slot, first_iter, n_iter, _fp(adj), _dp(gs),
                                     _dp(ge) if ge is not None else None, C.byref(ms)), "pg3g_adjoint")
        out = (gs[:self.num_scalars], ge)
        return out + (float(ms.value),) if timed else out


class DiffViews:
    """Several cameras of one DiffScene rendered and differentiated together
    (pg3g_diff_set_views): one launch for the forward of every view and one for the adjoint, with
    the gradients summed over the views.  `scenes`: SceneBuild objects of the same scene seen from
    other cameras; each view takes its camera and its RNG scene id (so it renders the samples that
    scene would render in a slot of its own)."""

    def __init__(self, diff_scene, scenes):
        self._lib = diff_scene._lib
        self.ds = diff_scene
        n = len(scenes)
        cams = (_ffi.CameraDesc * n)(*[sb.desc.camera for sb in scenes])
        ids = (C.c_uint32 * n)(*[int(sb.desc.params.scene_id) for sb in scenes])
        check(self._lib.pg3g_diff_set_views(diff_scene.ctx.ptr, diff_scene.slot, n, cams, ids), "pg3g_diff_set_views")
        self.num_views = n
        w, h = diff_scene.resolution
        self.shape = (n, h, w, 3)

    def render(self, first_iter, n_iter, timed=False):
        """Mean images (views, H, W, 3) FP64 of iterations [first_iter, first_iter + n_iter)."""
        out = np.zeros(self.shape, np.float64)
        ms = C.c_float()
        check(self._lib.pg3g_render_diff_views(self.ds.ctx.ptr, self.ds.slot, first_iter, n_iter, _dp(out),
                                               C.byref(ms) if timed else None), "pg3g_render_diff_views")
        out /= max(n_iter, 1)
        return (out, float(ms.value)) if timed else out

    def adjoint(self, first_iter, n_iter, dloss_dimages, timed=False):
        """Gradients (summed over the views) of a loss whose derivative with respect to the mean
        images of the same iterations is `dloss_dimages` (views, H, W, 3)."""
        adj = np.ascontiguousarray(dloss_dimages, np.float32).reshape(self.shape)
        ds = self.ds
        gs = np.zeros(max(ds.num_scalars, 1), np.float64)
        ge = np.zeros(ds.em_shape, np.float64) if ds.em_shape else None
        ms = C.c_float()
        check(self._lib.pg3g_adjoint_views(ds.ctx.ptr, ds.slot, first_iter, n_iter, _fp(adj), _dp(gs),
                                           _dp(ge) if ge is not None else None, C.byref(ms)), "pg3g_adjoint_views")
        out = (gs[:ds.num_scalars], ge)
        return out + (float(ms.value),) if timed else out


def render_function(diff_scene, n_iter, decorrelate=False):
    """A torch.autograd.Function for `diff_scene`: apply(scalars, em) -> image (H, W, 3) float64,
    or, for a DiffViews, the images of all its views (views, H, W, 3) from one launch.

    Each call renders the next `n_iter` iterations (fresh samples every call); the parameters of
    the call become the sampling parameters.  Backward replays exactly those samples, or, with
    `decorrelate`, the next `n_iter` iterations at the same parameters: the loss gradient then
    pairs the residual of one set of samples with the image derivative of an independent set, so
    for a loss like L2 (image - target)^2 its expectation is the gradient of the loss of the
    expected image, without the covariance of image and derivative that shared samples add."""
    import torch

    state = {"next": 0}
    span = 2 * n_iter if decorrelate else n_iter

    # a DiffViews renders and differentiates all its views at once; a DiffScene its own camera
    scene = diff_scene.ds if isinstance(diff_scene, DiffViews) else diff_scene

    class RenderFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, scalars, em):
            first = state["next"]
            state["next"] += span
            scene.set_params(scalars.detach().cpu().numpy(),
                             None if em is None else em.detach().cpu().numpy(), refresh_sampling=True)
            img = diff_scene.render(first, n_iter)
            ctx.first = first
            ctx.has_em = em is not None
            ctx.dtypes = (scalars.dtype, None if em is None else em.dtype)
            return torch.from_numpy(img)

        @staticmethod
        def backward(ctx, grad_img):
            gs, ge = diff_scene.adjoint(ctx.first + span - n_iter, n_iter, grad_img.detach().cpu().numpy())
            g_em = torch.from_numpy(ge).to(ctx.dtypes[1]) if (ctx.has_em and ge is not None) else None
            return torch.from_numpy(gs).to(ctx.dtypes[0]), g_em

    return RenderFn
