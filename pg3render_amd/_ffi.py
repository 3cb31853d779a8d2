"""ctypes view of the C ABI in include/pg3g.h and include/pg3g_scenes.h.

The structures mirror the header field by field.  `load()` opens the in-tree libpg3g.so and fails
loudly when it is missing: there is no CPU fallback for the GPU path.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpg3g.so")

PG3G_OK = 0
PG3G_ERR_INVALID_ARG = 1
PG3G_ERR_HIP = 2
PG3G_ERR_UNSUPPORTED = 3
PG3G_ERR_OUT_OF_MEMORY = 4
PG3G_ERR_NO_SCENE = 5

# Algorithm numbering of `enum Algorithm` (src/config.hxx:23-34)
ALG_DIRECT_BSDF = 2
ALG_DIRECT_LIGHT_ALL = 3
ALG_DIRECT_LIGHT_SINGLE = 4
ALG_DIRECT_MIS = 5
ALG_PATH_TRACING_NAIVE = 6
ALG_PATH_TRACING = 7
ALGORITHM_ACRONYMS = {"dbs": 2, "dlsa": 3, "dlss": 4, "dmis": 5, "ptn": 6, "pt": 7}

MAT_LAMBERT, MAT_PHONG, MAT_GGX_CONDUCTOR, MAT_GGX_DIELECTRIC, MAT_SMOOTH_CONDUCTOR, MAT_SMOOTH_DIELECTRIC = range(6)
LIGHT_AREA, LIGHT_POINT, LIGHT_BACKGROUND = range(3)
EM_NONE, EM_COSINE, EM_SIMPLE, EM_STEERABLE = range(4)
KERNEL_WAVEFRONT, KERNEL_MEGAKERNEL = 0, 1

F3 = C.c_float * 3


class Material(C.Structure):
    _fields_ = [("type", C.c_int32), ("light_id", C.c_int32), ("diffuse", F3), ("glossy", F3),
                ("phong_exponent", C.c_float), ("phong_diffuse_on", C.c_int32), ("phong_glossy_on", C.c_int32),
                ("alpha", C.c_float), ("inner_ior", C.c_float), ("outer_ior", C.c_float),
                ("absorbance", C.c_float), ("back_side_light_sampling", C.c_int32)]


class Light(C.Structure):
    _fields_ = [("type", C.c_int32), ("p0", F3), ("p1", F3), ("p2", F3), ("power", F3)]


class EmDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("texels", C.POINTER(C.c_float)),
                ("bilinear", C.c_int32), ("sampler", C.c_int32), ("st_max_approx_error", C.c_float),
                ("st_min_subdiv_level", C.c_int32), ("st_max_subdiv_level", C.c_int32)]


class CameraDesc(C.Structure):
    _fields_ = [("position", F3), ("forward", F3), ("up", F3), ("hfov_deg", C.c_float),
                ("res_x", C.c_uint32), ("res_y", C.c_uint32)]


class RenderParams(C.Structure):
    _fields_ = [("algorithm", C.c_int32), ("min_path_length", C.c_uint32), ("max_path_length", C.c_uint32),
                ("indirect_clipping", C.c_float), ("splitting_budget", C.c_float), ("splitting_level", C.c_float),
                ("light_to_bsdf_ratio", C.c_float), ("seed", C.c_uint64), ("scene_id", C.c_uint32)]


class SceneDesc(C.Structure):
    _fields_ = [("num_triangles", C.c_uint32), ("triangles", C.POINTER(C.c_float)),
                ("triangle_mat", C.POINTER(C.c_int32)), ("num_spheres", C.c_uint32),
                ("spheres", C.POINTER(C.c_float)), ("sphere_mat", C.POINTER(C.c_int32)),
                ("num_materials", C.c_uint32), ("materials", C.POINTER(Material)),
                ("num_lights", C.c_uint32), ("lights", C.POINTER(Light)), ("em", EmDesc),
                ("camera", CameraDesc), ("params", RenderParams)]


class SteerableBuildInfo(C.Structure):
    _fields_ = [("num_leaves", C.c_uint32), ("num_vertices", C.c_uint32), ("num_levels", C.c_uint32),
                ("threads", C.c_uint32), ("from_cache", C.c_uint32), ("build_ms", C.c_double)]


class ImageDiff(C.Structure):
    _fields_ = [("pixel_rel_rmse", C.c_double), ("image_rel_rmse", C.c_double), ("max_pixel_rel", C.c_double),
                ("mean_luminance", C.c_double), ("mean_luminance_reference", C.c_double)]


class BatchConfig(C.Structure):
    _fields_ = [("iters_per_item", C.c_uint32), ("waves_per_cu", C.c_uint32), ("use_graph", C.c_uint32)]


_P = C.c_void_p
_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int32)
_DP = C.POINTER(C.c_double)
_U64P = C.POINTER(C.c_uint64)
_U32P = C.POINTER(C.c_uint32)

_SIGNATURES = {
    "pg3g_create": (C.c_int, [C.POINTER(_P), C.c_int]),
    "pg3g_destroy": (C.c_int, [_P]),
    "pg3g_last_error": (C.c_char_p, []),
    "pg3g_upload_scene": (C.c_int, [_P, C.POINTER(SceneDesc)]),
    "pg3g_upload_scene_slot": (C.c_int, [_P, C.c_uint32, C.POINTER(SceneDesc)]),
    "pg3g_set_params": (C.c_int, [_P, C.c_uint32, C.POINTER(RenderParams)]),
    "pg3g_render": (C.c_int, [_P, C.c_uint32, C.c_uint32]),
    "pg3g_render_slot": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32]),
    "pg3g_set_kernel_mode": (C.c_int, [_P, C.c_int]),
    "pg3g_render_timed": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, _FP]),
    "pg3g_read_framebuffer": (C.c_int, [_P, _DP, _U64P]),
    "pg3g_read_framebuffer_slot": (C.c_int, [_P, C.c_uint32, _DP, _U64P]),
    "pg3g_clear_framebuffer": (C.c_int, [_P, C.c_uint32]),
    "pg3g_framebuffer_device_ptr": (C.c_int, [_P, C.c_uint32, C.POINTER(_DP)]),
    "pg3g_synchronize": (C.c_int, [_P]),
    "pg3g_read_counters": (C.c_int, [_P, _U64P, _U64P]),
    "pg3g_read_ray_counters": (C.c_int, [_P, _U64P, _U64P]),
    "pg3g_run_iteration": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_int]),
    "pg3g_scene_primitives": (C.c_int, [_P, C.c_uint32, _U32P, _U32P, _U32P]),
    "pg3g_adjoint_noalloc": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pg3g_reset_counters": (C.c_int, [_P]),
    "pg3g_bsdf_eval_batch": (C.c_int, [_P, C.POINTER(Material), C.c_uint32, _FP, _FP, _FP, _FP, _FP]),
    "pg3g_bsdf_sample_batch": (C.c_int, [_P, C.POINTER(Material), C.c_uint32, _FP, _FP, _FP, _FP, _FP, _FP]),
    "pg3g_em_sample_batch": (C.c_int, [_P, C.c_uint32, _FP, _IP, _FP, _FP, _FP, _FP, _IP]),
    "pg3g_em_pdf_batch": (C.c_int, [_P, C.c_uint32, _FP, _FP, _IP, _FP]),
    "pg3g_em_eval_batch": (C.c_int, [_P, C.c_uint32, _FP, _FP]),
    "pg3g_intersect_batch": (C.c_int, [_P, C.c_uint32, _FP, _FP, _FP, _FP, _FP, _IP]),
    "pg3g_occluded_batch": (C.c_int, [_P, C.c_uint32, _FP, _FP, _FP, _IP]),
    "pg3g_steerable_info": (C.c_int, [_P, _U32P, _U32P, _U32P]),
    "pg3g_steerable_leaves": (C.c_int, [_P, _FP]),
    "pg3g_steerable_root_weight": (C.c_int, [_P, _FP]),
    "pg3g_steerable_build": (C.c_int, [C.POINTER(EmDesc), C.c_uint32, C.POINTER(SteerableBuildInfo), _FP, C.c_uint32,
                                       _FP]),
    "pg3g_steerable_cache_dir": (C.c_int, [C.c_char_p]),
    "pg3g_load_image": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), _FP, C.c_uint64]),
    "pg3g_load_em": (C.c_int, [C.c_char_p, C.c_float, C.c_float, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), _FP,
                               C.c_uint64]),
    "pg3g_save_image": (C.c_int, [C.c_char_p, _DP, C.c_uint32, C.c_uint32, C.c_double, C.c_float]),
    "pg3g_image_compare": (C.c_int, [_FP, _FP, C.c_uint32, C.c_uint32, C.POINTER(ImageDiff), _FP]),
    "pg3g_batch_create": (C.c_int, [_P, C.c_uint32, _U32P, _U32P, C.POINTER(_P)]),
    "pg3g_batch_destroy": (C.c_int, [_P]),
    "pg3g_batch_set_config": (C.c_int, [_P, C.POINTER(BatchConfig)]),
    "pg3g_batch_get_config": (C.c_int, [_P, C.POINTER(BatchConfig)]),
    "pg3g_batch_render": (C.c_int, [_P, C.c_uint32]),
    "pg3g_batch_render_timed": (C.c_int, [_P, C.c_uint32, _FP]),
    "pg3g_batch_autotune": (C.c_int, [_P, C.c_uint32, _FP]),
    "pg3g_diff_setup": (C.c_int, [_P, C.c_uint32, C.c_uint32, _U32P, _U32P]),
    "pg3g_diff_get_params": (C.c_int, [_P, C.c_uint32, _FP, _FP]),
    "pg3g_diff_set_params": (C.c_int, [_P, C.c_uint32, _FP, _FP, C.c_int]),
    "pg3g_render_diff": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, _FP]),
    "pg3g_adjoint": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, _FP, _DP, _DP, _FP]),
    "pg3g_diff_set_views": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.POINTER(CameraDesc), _U32P]),
    "pg3g_render_diff_views": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, _DP, _FP]),
    "pg3g_adjoint_views": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, _FP, _DP, _DP, _FP]),
    "pg3g_build_workload_scene": (C.c_int, [C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(_P)]),
    "pg3g_build_reference_scene": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(_P)]),
    "pg3g_build_reference_scene_em": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_char_p, C.POINTER(_P)]),
    "pg3g_reference_scene_acronym": (C.c_int, [C.c_uint32, C.c_uint32, C.c_char_p, C.c_size_t]),
    "pg3g_scene_build_desc": (C.POINTER(SceneDesc), [_P]),
    "pg3g_scene_build_mutable": (C.POINTER(SceneDesc), [_P]),
    "pg3g_scene_build_em_texels": (_FP, [_P]),
    "pg3g_scene_build_spp": (C.c_uint32, [_P]),
    "pg3g_scene_build_free": (None, [_P]),
}

# Entry points declared in include/*.h, used by the export test.
EXPORTED_SYMBOLS = sorted(_SIGNATURES)

_lib = None


class Pg3gError(RuntimeError):
    def __init__(self, code, where, msg):
        super().__init__(f"{where} failed with status {code}: {msg}")
        self.code = code


def load(path=LIB_PATH):
    """Opens libpg3g.so (built in-tree by __graft_entry__.build()) and declares its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libpg3g.so is not built ({path}); run __graft_entry__.build() first")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code, where):
    if code != PG3G_OK:
        msg = _lib.pg3g_last_error().decode() if _lib is not None else ""
        raise Pg3gError(code, where, msg)
