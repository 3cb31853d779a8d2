"""Inverse rendering with the GPU adjoint: recover an environment map and sphere roughnesses.

Workload: a floor with four GGX-conductor spheres (seeded roughness alpha in U[0.05, 0.5]) lit by
a 128x64 RGB EM of three Gaussian lobes on a sky gradient, seen from 8 cameras circling the
spheres (pg3g_build_workload_scene(5, seed, view)).  Targets are rendered from the hidden ground
truth; the optimisation starts from alpha = 0.3 and a constant grey EM at the ground truth's mean
radiance, and runs Adam (FP32 parameters) on the logs of alpha and of the EM texels (both stay
positive and move by comparable relative amounts per step), each step
rendering every view with fresh samples and back-propagating through the path-replay adjoint
(FP64 EM-gradient accumulation on the GPU), via the torch.autograd.Function of each view.
"""
import time

import numpy as np

from .diff import DiffScene, render_function
from .renderer import SceneBuild

SPHERE_MATS = [3, 4, 5, 6]      # workload 5: floor materials 0..2, then the four spheres


def inverse_rendering(ctx, steps=200, res=256, spp=16, target_spp=64, lr=1e-2, seed=2024, views=8, log_every=0):
    import torch

    scenes = [SceneBuild.workload(5, seed=seed, variant=v, res=res) for v in range(views)]
    diffs = []
    for v, sb in enumerate(scenes):
        ctx.upload(sb, v)
        diffs.append(DiffScene(ctx, sb, SPHERE_MATS, slot=v))
    theta_gt, em_gt = diffs[0].get_params()
    alpha_gt = theta_gt[3::4].copy()
    targets = [torch.from_numpy(d.render(1 << 20, target_spp)) for d in diffs]

    tints = torch.tensor(theta_gt.reshape(4, 4)[:, :3])             # known, not optimised
    alpha_log = torch.full((4,), float(np.log(0.3)), dtype=torch.float32, requires_grad=True)
    em_log = torch.full(em_gt.shape, float(np.log(max(em_gt.mean(), 1e-3))), dtype=torch.float32,
                        requires_grad=True)
    opt = torch.optim.Adam([alpha_log, em_log], lr=lr)
    fns = [render_function(d, spp) for d in diffs]

    def errors():
        a = np.exp(alpha_log.detach().numpy())
        em = np.exp(em_log.detach().numpy())
        return (float(np.abs(a - alpha_gt).mean()),
                float(np.sqrt(np.mean((em - em_gt) ** 2)) / np.sqrt(np.mean(em_gt ** 2))))

    a0, e0 = errors()
    history = []
    t0 = time.perf_counter()
    for step in range(steps):
        opt.zero_grad()
        scal = torch.cat([tints, torch.exp(alpha_log)[:, None]], dim=1).reshape(-1)
        em = torch.exp(em_log)
        loss = 0.0
        for fn, tgt in zip(fns, targets):
            img = fn.apply(scal, em)
            loss = loss + torch.mean((img - tgt) ** 2) / len(fns)
        loss.backward()
        opt.step()
        history.append(float(loss.detach()))
        if log_every and step % log_every == 0:
            print(f"step {step}: loss {history[-1]:.6g}, alpha {np.exp(alpha_log.detach().numpy())}", flush=True)
    wall = time.perf_counter() - t0
    a1, e1 = errors()
    return {"steps": steps, "views": views, "resolution": res, "spp": spp, "loss_first": history[0],
            "loss_last": history[-1], "loss_history": history, "alpha_gt": alpha_gt.tolist(),
            "alpha_final": np.exp(alpha_log.detach().numpy()).tolist(), "alpha_abs_err_start": a0, "alpha_abs_err_end": a1,
            "em_rel_rmse_start": e0, "em_rel_rmse_end": e1, "ms_per_step": wall * 1e3 / max(steps, 1)}
