"""C-ABI library checks that need no GPU: libpg3g.so loads and exports every entry point that
include/*.h declares; the scene builders produce valid descriptions; errors come back as status
codes with messages (no exceptions across the ABI)."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from pg3render_amd import SceneBuild, _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"^[\w\s\*]*?\b(pg3g_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_ffi.LIB_PATH)
    decl = declared_functions()
    assert len(decl) >= 30
    missing = [n for n in decl if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding declares signatures for every header entry point
    assert set(decl) == set(_ffi.EXPORTED_SYMBOLS), set(decl) ^ set(_ffi.EXPORTED_SYMBOLS)


@pytest.mark.parametrize("workload,variant", [(1, 0), (2, 0), (3, 5), (4, 0), (5, 3)])
def test_workload_scene_builders(workload, variant):
    sb = SceneBuild.workload(workload, seed=11, variant=variant)
    d = sb.desc
    assert d.num_materials > 0 and d.camera.res_x > 0
    tris = np.ctypeslib.as_array(d.triangles, shape=(d.num_triangles, 9)) if d.num_triangles else np.zeros((0, 9))
    mats = np.ctypeslib.as_array(d.triangle_mat, shape=(d.num_triangles,)) if d.num_triangles else []
    assert np.all(np.asarray(mats) < d.num_materials)
    assert np.all(np.isfinite(tris))
    if workload in (2, 3, 4, 5):
        em = sb.em_texels()
        assert em is not None and np.all(em >= 0) and np.isfinite(em).all()
    assert sb.spp in (4, 16, 32, 64, 256)


def test_workload5_views_share_one_scene():
    """The 8 views of the inverse-rendering workload see one scene: the sphere materials and the
    EM depend on the seed only; the cameras differ."""
    def content(v):
        sb = SceneBuild.workload(5, seed=2024, variant=v, res=32)
        d = sb.desc
        mats = [(m.type, tuple(m.glossy), m.alpha, tuple(m.diffuse)) for m in (d.materials[i] for i in range(d.num_materials))]
        sph = np.ctypeslib.as_array(d.spheres, shape=(d.num_spheres, 4)).copy()
        return mats, sb.em_texels().copy(), sph, tuple(d.camera.position)

    m0, e0, s0, c0 = content(0)
    alphas = [m[2] for m in m0 if m[0] == _ffi.MAT_GGX_CONDUCTOR]
    assert len(alphas) == 4 and len(set(alphas)) == 4
    for v in range(1, 8):
        m, e, s, c = content(v)
        assert m == m0 and np.array_equal(e, e0) and np.array_equal(s, s0)
        assert c != c0
    # another seed is another scene
    a = SceneBuild.workload(5, seed=2024, variant=0, res=32).em_texels().copy()
    b = SceneBuild.workload(5, seed=2025, variant=0, res=32).em_texels().copy()
    assert not np.array_equal(a, b)


def test_workload1_light_is_radiance_25():
    sb = SceneBuild.workload(1)
    d = sb.desc
    l0 = d.lights[0]
    p = [np.array(l0.p0), np.array(l0.p1), np.array(l0.p2)]
    area = 0.5 * np.linalg.norm(np.cross(p[1] - p[0], p[2] - p[0]))
    assert abs(l0.power[0] / (np.pi * area) - 25.0) < 1e-3
    n = np.cross(p[1] - p[0], p[2] - p[0])
    assert n[2] < 0          # emits downwards into the box


def test_reference_scene_ids_build():
    for sid in list(range(0, 27)) + list(range(28, 38)):
        sb = SceneBuild.reference(sid, em_type=0, resolution=32)
        assert sb.desc.num_materials == 11
    with pytest.raises(_ffi.Pg3gError) as e:
        SceneBuild.reference(27)       # SRA layered material is outside the GPU path
    assert e.value.code == _ffi.PG3G_ERR_UNSUPPORTED


def test_errors_are_status_codes():
    lib = _ffi.load()
    assert lib.pg3g_render(None, 0, 1) == _ffi.PG3G_ERR_NO_SCENE
    assert lib.pg3g_synchronize(None) == _ffi.PG3G_ERR_INVALID_ARG
    assert b"null" in lib.pg3g_last_error()
