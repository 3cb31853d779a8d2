"""pg3gpu renders on the GPU: the image it writes for a reference scene is the library's image of
the same scene and settings (up to RGBE rounding), and the time-based mode (-t, the reference's
pg3render.cxx:73-100 loop) renders at least one iteration and writes the default-named file."""
import os
import subprocess

import numpy as np
import pytest

from pg3render_amd import Context, SceneBuild, _ffi
from pg3render_amd.renderer import load_image

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PG3GPU = os.path.join(ROOT, "pg3render_amd", "pg3gpu")

pytestmark = pytest.mark.gpu


def run(*args):
    assert os.path.exists(PG3GPU), "pg3gpu is not built (make -C pg3render_amd/csrc)"
    print("pg3gpu", *args, flush=True)
    return subprocess.run([PG3GPU, *args], capture_output=True, text=True, timeout=100)


def test_cli_image_is_the_library_image(tmp_path):
    r = run("-s", "2", "-r", "32", "-i", "8", "-maxpl", "6", "-sb", "1", "-iic", "0", "-seed", "7",
            "-k", "megakernel", "-od", str(tmp_path), "-o", "a.hdr", "-q")
    assert r.returncode == 0, r.stdout + r.stderr
    cli = load_image(str(tmp_path / "a.hdr")).astype(np.float64)

    ctx = Context()
    try:
        ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)
        sb = SceneBuild.reference(2, em_type=0, resolution=32)
        sb.set_params(algorithm="pt", min_path_length=1, max_path_length=6, splitting_budget=1.0,
                      splitting_level=1.0, light_to_bsdf_ratio=1.0, indirect_clipping=0.0, seed=7)
        ctx.upload(sb)
        ctx.render(0, 8)
        ctx.synchronize()
        fb, it = ctx.framebuffer(0, sb.resolution)
    finally:
        ctx.close()
    assert it == 8
    lib = fb / it
    assert cli.shape == lib.shape and lib.mean() > 0
    # RGBE keeps 8 mantissa bits relative to a pixel's largest channel
    tol = lib.max(axis=2, keepdims=True) * 2.0 ** -7 + 1e-6
    assert np.all(np.abs(cli - lib) <= tol), np.abs(cli - lib).max()


def test_cli_time_based_render(tmp_path):
    name = run("-s", "2", "-r", "32", "-t", "0.5", "-opof").stdout
    assert name == "w2s_c_SpdWpd_pt_rr_splt4,1.0,1.0_iic16.0_0.5sec.bmp", name
    r = run("-s", "2", "-r", "32", "-t", "0.5", "-od", str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    assert " iterations," in r.stdout
    iters = int(r.stdout.split(" iterations,")[0].split()[-1])
    assert iters >= 1
    assert (tmp_path / name).stat().st_size == 54 + 32 * 32 * 3
