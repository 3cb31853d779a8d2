"""Inverse rendering through the torch wrapper of the adjoint, at a reduced size: the loss and the
roughness error must fall over a few dozen Adam steps."""
import numpy as np
import pytest

from pg3render_amd import Context
from pg3render_amd.inverse import inverse_rendering

pytestmark = pytest.mark.gpu


def test_inverse_rendering_reduces_loss_and_roughness_error():
    ctx = Context(0)
    try:
        r = inverse_rendering(ctx, steps=30, res=48, spp=4, target_spp=32, views=4)
    finally:
        ctx.close()
    print(f"loss {r['loss_first']:.4g} -> {r['loss_last']:.4g}, alpha error {r['alpha_abs_err_start']:.4f} -> "
          f"{r['alpha_abs_err_end']:.4f}, loss floor {r['loss_floor']:.4g}, EM rel RMSE {r['em_rel_rmse_start']:.4f} -> {r['em_rel_rmse_end']:.4f}, "
          f"{r['ms_per_step']:.1f} ms/step")
    assert np.all(np.isfinite(r["loss_history"]))
    assert np.mean(r["loss_history"][-5:]) < 0.8 * np.mean(r["loss_history"][:3])
    assert r["alpha_abs_err_end"] < r["alpha_abs_err_start"]
    assert r["em_upper_rel_rmse_end"] < r["em_upper_rel_rmse_start"]
    assert np.isfinite(r["em_rel_rmse_end"]) and r["em_rel_rmse_end"] < r["em_rel_rmse_start"]
