"""Inverse rendering with the GPU adjoint: recover an environment map and sphere roughnesses.

Workload: a floor with four GGX-conductor spheres (seeded roughness alpha in U[0.05, 0.5]) lit by
a 128x64 RGB EM of three Gaussian lobes on a sky gradient, seen from 8 cameras circling the
spheres (pg3g_build_workload_scene(5, seed, view)).  Targets are rendered from the hidden ground
truth; the optimisation starts from alpha = 0.3 and a constant grey EM at the ground truth's mean
radiance, and runs Adam (FP32 parameters) on the logs of alpha and of the EM texels (both stay
positive and move by comparable relative amounts per step), each step rendering every view with
fresh samples and back-propagating through the path-replay adjoint (FP64 EM-gradient accumulation
on the GPU).  The 8 views are one slot with 8 cameras (DiffViews): each step renders all of them in
one launch and back-propagates through one adjoint launch, via one torch.autograd.Function.

Defaults (chosen by a sweep on this workload): steps of 0.02 (alpha) and 0.1 (EM) under a cosine
decay to a tenth, a smoothness prior of weight 0.1 on the log EM, and gradients from samples
independent of the rendered image.  With the image's own samples the gradient of the L2 loss
carries the covariance of image and derivative, which biased every roughness low (0.08 against
0.15); with independent samples the 200 steps recover the four roughnesses to 0.004 and the
visible half of the EM to about 0.1 relative RMSE, with the loss at its noise floor (the ground
truth rendered at the same spp, `loss_floor`).
"""
import time

import numpy as np

from .diff import DiffScene, DiffViews, render_function
from .renderer import SceneBuild

SPHERE_MATS = [3, 4, 5, 6]      # workload 5: floor materials 0..2, then the four spheres


def inverse_rendering(ctx, steps=200, res=256, spp=16, target_spp=64, lr=0.02, lr_em=0.1, seed=2024, views=8,
                      log_every=0, decay=0.1, decorrelate=True, em_smooth=0.1):
    """lr: Adam step of log alpha; lr_em: of the log EM texels (None: lr); decay: final fraction
    of both steps under a cosine schedule (0: constant steps); decorrelate: gradients from
    samples independent of the rendered image (render_function); em_smooth: weight of a
    smoothness prior on the log EM (mean squared difference of neighbouring texels, longitude
    wrapping).  Texels that few paths reach get rare, noisy gradients, which Adam's per-parameter
    normalisation turns into full-size steps; the prior ties them to their neighbours."""
    import torch

    scenes = [SceneBuild.workload(5, seed=seed, variant=v, res=res) for v in range(views)]
    ctx.upload(scenes[0], 0)
    diff = DiffScene(ctx, scenes[0], SPHERE_MATS, slot=0)
    multi = DiffViews(diff, scenes)       # view v: the camera and RNG streams of scenes[v]
    theta_gt, em_gt = diff.get_params()
    alpha_gt = theta_gt[3::4].copy()
    targets = torch.from_numpy(multi.render(1 << 20, target_spp))
    # noise floor of the loss: the ground truth rendered at the optimisation's spp
    loss_floor = float(np.mean((multi.render(1 << 21, spp) - targets.numpy()) ** 2))

    tints = torch.tensor(theta_gt.reshape(4, 4)[:, :3])             # known, not optimised
    alpha_log = torch.full((4,), float(np.log(0.3)), dtype=torch.float32, requires_grad=True)
    em_log = torch.full(em_gt.shape, float(np.log(max(em_gt.mean(), 1e-3))), dtype=torch.float32,
                        requires_grad=True)
    opt = torch.optim.Adam([{"params": [alpha_log], "lr": lr}, {"params": [em_log], "lr": lr if lr_em is None else lr_em}])
    base_lrs = [g["lr"] for g in opt.param_groups]
    # texels of the upper hemisphere (lat-long rows with v < 1/2): the floor hides the rest from
    # every view, so only these are recoverable
    upper = np.zeros(em_gt.shape[:2], bool)
    upper[: em_gt.shape[0] // 2] = True
    fn = render_function(multi, spp, decorrelate)

    def rel_rmse(em, mask=None):
        d, g = (em - em_gt), em_gt
        if mask is not None:
            d, g = d[mask], g[mask]
        return float(np.sqrt(np.mean(d ** 2)) / np.sqrt(np.mean(g ** 2)))

    def errors():
        a = np.exp(alpha_log.detach().numpy())
        em = np.exp(em_log.detach().numpy())
        return float(np.abs(a - alpha_gt).mean()), rel_rmse(em), rel_rmse(em, upper)

    a0, e0, u0 = errors()
    history = []
    t0 = time.perf_counter()
    for step in range(steps):
        if decay > 0.0:
            f = decay + (1.0 - decay) * 0.5 * (1.0 + np.cos(np.pi * step / max(steps - 1, 1)))
            for g, b in zip(opt.param_groups, base_lrs):
                g["lr"] = b * f
        opt.zero_grad()
        scal = torch.cat([tints, torch.exp(alpha_log)[:, None]], dim=1).reshape(-1)
        em = torch.exp(em_log)
        imgs = fn.apply(scal, em)
        loss = torch.mean((imgs - targets) ** 2)     # the mean over the views of each view's L2 loss
        total = loss
        if em_smooth > 0.0:
            dx = em_log - torch.roll(em_log, 1, dims=1)
            dy = em_log[1:] - em_log[:-1]
            total = total + em_smooth * (torch.mean(dx ** 2) + torch.mean(dy ** 2))
        total.backward()
        opt.step()
        history.append(float(loss.detach()))
        if log_every and step % log_every == 0:
            print(f"step {step}: loss {history[-1]:.6g}, alpha {np.exp(alpha_log.detach().numpy())}", flush=True)
    wall = time.perf_counter() - t0
    a1, e1, u1 = errors()
    return {"steps": steps, "views": views, "resolution": res, "spp": spp, "loss_first": history[0],
            "loss_last": history[-1], "loss_floor": loss_floor, "loss_history": history, "alpha_gt": alpha_gt.tolist(),
            "alpha_final": np.exp(alpha_log.detach().numpy()).tolist(), "alpha_abs_err_start": a0, "alpha_abs_err_end": a1,
            "em_rel_rmse_start": e0, "em_rel_rmse_end": e1, "em_upper_rel_rmse_start": u0, "em_upper_rel_rmse_end": u1,
            "ms_per_step": wall * 1e3 / max(steps, 1)}
