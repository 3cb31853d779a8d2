"""Inverse rendering through the torch wrapper of the adjoint at the workload's own size: 200 Adam
steps over 8 views of 256x256 pixels at 16 spp (max path length 6), recovering a 128x64 RGB EM and
the roughness of four GGX spheres.  The loss must reach its noise floor (the ground truth rendered
at the same spp against the same targets) and the roughness error must fall below 0.01."""
import numpy as np
import pytest

from pg3render_amd import Context
from pg3render_amd.inverse import inverse_rendering

pytestmark = pytest.mark.gpu


def test_inverse_rendering_recovers_roughness_and_em():
    ctx = Context(0)
    try:
        r = inverse_rendering(ctx, steps=200, res=256, spp=16, views=8)
    finally:
        ctx.close()
    hist = np.asarray(r["loss_history"])
    print(f"loss {r['loss_first']:.4g} -> {r['loss_last']:.4g} (mean of the last 10 {hist[-10:].mean():.4g}), "
          f"loss floor {r['loss_floor']:.4g}; alpha error {r['alpha_abs_err_start']:.4f} -> {r['alpha_abs_err_end']:.4f} "
          f"(ground truth {np.round(r['alpha_gt'], 4)}, final {np.round(r['alpha_final'], 4)}); EM rel RMSE "
          f"{r['em_rel_rmse_start']:.4f} -> {r['em_rel_rmse_end']:.4f} (above the horizon "
          f"{r['em_upper_rel_rmse_start']:.4f} -> {r['em_upper_rel_rmse_end']:.4f}); {r['ms_per_step']:.1f} ms/step")
    assert r["steps"] == 200 and r["views"] == 8 and r["resolution"] == 256 and r["spp"] == 16
    assert np.all(np.isfinite(hist))
    # the loss ends at its noise floor: what is left is the sampling noise of the 16-spp images
    assert hist[-10:].mean() <= 1.05 * r["loss_floor"], (hist[-10:].mean(), r["loss_floor"])
    assert r["alpha_abs_err_end"] < 0.01
    assert r["em_upper_rel_rmse_end"] < 0.5 * r["em_upper_rel_rmse_start"]
    assert r["em_rel_rmse_end"] < r["em_rel_rmse_start"]
