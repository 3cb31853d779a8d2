"""MI355X (gfx950) path-tracing hot path of PG3Render behind its renderer boundary.

The GPU work runs in libpg3g.so (hand-written HIP kernels + C ABI, include/pg3g.h); this package
only binds it.  There is no CPU fallback: if the library is not built, importing the renderer
raises.
"""
from . import _ffi
from .diff import DiffScene, DiffViews, render_function
from .renderer import (Batch, Context, GpuPathTracer, SceneBuild, compare_images, load_image, save_image,
                       steerable_build, steerable_cache_dir)

__all__ = ["Batch", "Context", "DiffScene", "DiffViews", "GpuPathTracer", "SceneBuild", "_ffi", "compare_images", "load_image",
           "render_function", "save_image", "steerable_build", "steerable_cache_dir"]
