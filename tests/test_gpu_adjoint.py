"""Path-replay adjoint (pg3g_adjoint) against central finite differences.

Scene: the differentiation workload at its own size (box, two GGX-conductor spheres with RGB tint
and roughness, 128x64 sky EM sampled by the simple spherical sampler, 256x256 pixels, 32 spp).
Loss: mean squared error against a target rendered from hidden parameters.  Finite differences use
the same samples (correlated RNG: identical iteration range) and frozen sampling, with the loss
accumulated in FP64 on the CPU, h = 1e-3.  Parameters checked: all 11 scalars (two conductor tints
and roughnesses, the white wall's albedo) plus texels drawn uniformly at random -- 53 from all
24,576 EM texel channels and 64 from the rows above the horizon (the part of the EM the box's open
side lets the camera paths reach) -- chosen without looking at the adjoint's output, so a texel
whose gradient the adjoint drops is as likely to be drawn as any other.  Required agreement: 2 %.
"""
import numpy as np
import pytest

from pg3render_amd import Context, SceneBuild
from pg3render_amd.diff import DiffScene, render_function

pytestmark = pytest.mark.gpu

RES, SPP, H = 256, 32, 1e-3
WHITE, COND0, COND1 = 0, 3, 4      # material ids of workload 4: walls 0..2, conductors 3, 4


@pytest.fixture(scope="module")
def setup():
    ctx = Context(0)
    sb = SceneBuild.workload(4, seed=2024, res=RES)
    ctx.upload(sb, 0)
    ds = DiffScene(ctx, sb, [COND0, COND1, WHITE], slot=0)
    theta0, em0 = ds.get_params()
    rng = np.random.default_rng(11)
    hidden = theta0 * rng.uniform(0.7, 1.3, theta0.shape).astype(np.float32)
    em_hidden = em0 * rng.uniform(0.5, 1.5, em0.shape).astype(np.float32)
    ds.set_params(hidden, em_hidden, refresh_sampling=True)
    target = ds.render(1000, 64)
    ds.set_params(theta0, em0, refresh_sampling=True)
    yield ctx, sb, ds, theta0, em0, target
    ctx.close()


def loss(img, target):
    return float(np.mean((img - target) ** 2))


def test_scalar_layout(setup):
    _, _, ds, theta0, em0, _ = setup
    assert ds.num_scalars == 11
    # conductor 0: tint RGB, alpha; conductor 1: tint RGB, alpha; white wall: albedo RGB
    assert 0.05 <= theta0[3] <= 0.5 and 0.05 <= theta0[7] <= 0.5
    np.testing.assert_allclose(theta0[8:11], [0.8, 0.8, 0.8], rtol=1e-6)
    assert em0.shape == (64, 128, 3)


def test_frozen_sampling_forward_matches_plain_render(setup):
    """At the sampling parameters the differentiable forward estimates the same image as the
    plain path tracer (same samples; values differ only by the BSDF evaluation rounding)."""
    ctx, sb, ds, *_ = setup
    a = ds.render(0, SPP)
    ctx.clear(0)
    ctx.render(0, SPP, 0)
    b, it = ctx.framebuffer(0, sb.resolution)
    b = b / it
    assert abs(a.sum() - b.sum()) / b.sum() < 1e-4
    assert np.mean(np.abs(a - b) > 1e-4 * np.abs(b) + 1e-6) < 0.01


def central_differences(ds, theta0, em0, target, params):
    """Central differences (h = H) of the FP64 loss over the same samples with sampling frozen."""
    out = []
    for kind, i in params:
        ls = []
        for sgn in (1.0, -1.0):
            th, em = theta0.copy(), em0.copy()
            if kind == "s":
                th[i] += sgn * H
            else:
                em.reshape(-1)[i] += sgn * H
            ds.set_params(th, em, refresh_sampling=False)
            ls.append(loss(ds.render(0, SPP), target))
        out.append((ls[0] - ls[1]) / (2 * H))
    ds.set_params(theta0, em0, refresh_sampling=True)
    return np.array(out)


def check_group(name, params, g, fd):
    """Relative error per parameter, floored at 1e-3 of the group's largest gradient (a texel
    whose FD and adjoint are both ~0 passes; one where only one of them is ~0 does not)."""
    scale = np.maximum(np.abs(fd), np.abs(g))
    floor = 1e-3 * max(scale.max(), 1e-30)
    rel = np.abs(g - fd) / np.maximum(scale, floor)
    worst = int(np.argmax(rel))
    zero = int(np.sum(np.abs(fd) < floor))
    print(f"{name}: {len(params)} parameters ({zero} with |FD| below the floor), median relative error "
          f"{np.median(rel):.2e}, max {rel.max():.2e} at {params[worst]} (adjoint {g[worst]:.4e}, FD {fd[worst]:.4e})")
    assert rel.max() <= 0.02, (name, params[worst], g[worst], fd[worst])


def test_adjoint_matches_central_differences(setup):
    _, _, ds, theta0, em0, target = setup
    img = ds.render(0, SPP)
    dl = 2.0 * (img - target) / img.size
    gs, ge = ds.adjoint(0, SPP, dl)
    ge = ge.reshape(-1)
    h, w, _ = em0.shape
    rng = np.random.default_rng(7)
    uniform = rng.choice(em0.size, 53, replace=False)
    upper = rng.choice((h // 2) * w * 3, 64, replace=False)       # rows 0 .. h/2-1 come first
    scal = [("s", i) for i in range(11)]
    tex_u = [("e", int(t)) for t in uniform]
    tex_h = [("e", int(t)) for t in upper]
    fd = central_differences(ds, theta0, em0, target, scal + tex_u + tex_h)
    check_group("scalars", scal, gs, fd[:11])
    check_group("texels, uniform over the EM", tex_u, ge[uniform], fd[11:64])
    check_group("texels, uniform above the horizon", tex_h, ge[upper], fd[64:])
    # the texels above the horizon carry gradient: the check is not vacuous
    assert np.sum(np.abs(fd[64:]) > 0) >= 32


def test_autograd_function_and_adjoint_cost(setup):
    import torch
    _, _, ds, theta0, em0, target = setup
    fn = render_function(ds, SPP)
    th = torch.tensor(theta0, requires_grad=True)
    em = torch.tensor(em0, requires_grad=True)
    img = fn.apply(th, em)
    l = torch.mean((img - torch.from_numpy(target)) ** 2)
    l.backward()
    first = 0   # the first call of a fresh render_function renders iterations [0, SPP)
    dl = 2.0 * (img.detach().numpy() - target) / img.numel()
    gs, ge = ds.adjoint(first, SPP, dl)
    np.testing.assert_allclose(th.grad.numpy(), gs, rtol=1e-4, atol=1e-9 * np.abs(gs).max())
    np.testing.assert_allclose(em.grad.numpy(), ge, rtol=1e-4, atol=1e-6 * np.abs(ge).max())
    _, fwd_ms = ds.render(0, SPP, timed=True)
    *_, adj_ms = ds.adjoint(0, SPP, dl, timed=True)
    print(f"forward {fwd_ms:.3f} ms, adjoint {adj_ms:.3f} ms, ratio {adj_ms / fwd_ms:.2f}")


def test_decorrelated_autograd_function(setup):
    """With decorrelate=True the image of call k is iterations [2k SPP, (2k+1) SPP) and its
    gradient is the adjoint of the next SPP iterations, weighted by that image's residual."""
    import torch
    _, _, ds, theta0, em0, target = setup
    ds.set_params(theta0, em0, refresh_sampling=True)
    fn = render_function(ds, SPP, decorrelate=True)
    for call in range(2):
        th = torch.tensor(theta0, requires_grad=True)
        em = torch.tensor(em0, requires_grad=True)
        img = fn.apply(th, em)
        np.testing.assert_array_equal(img.detach().numpy(), ds.render(2 * call * SPP, SPP))
        l = torch.mean((img - torch.from_numpy(target)) ** 2)
        l.backward()
        dl = 2.0 * (img.detach().numpy() - target) / img.numel()
        gs, ge = ds.adjoint((2 * call + 1) * SPP, SPP, dl)
        np.testing.assert_allclose(th.grad.numpy(), gs, rtol=1e-4, atol=1e-9 * np.abs(gs).max())
        np.testing.assert_allclose(em.grad.numpy(), ge, rtol=1e-4, atol=1e-6 * np.abs(ge).max())
        # an independent block: not the gradient of the image's own samples
        gs_own, _ = ds.adjoint(2 * call * SPP, SPP, dl)
        assert not np.allclose(gs_own, gs, rtol=1e-3)


def test_multi_view_forward_and_adjoint_equal_per_view_slots():
    """Eight cameras of the inverse-rendering workload as one slot (pg3g_diff_set_views): one
    launch renders every view and one adjoint launch returns the gradients summed over the views.
    View v takes the camera and RNG scene id of the scene built for view v, so its image equals
    that scene's own slot render (up to the order of FP64 additions), and the summed gradients
    equal the sum of the eight per-slot adjoints of the same samples."""
    from pg3render_amd import DiffViews
    ctx = Context(0)
    try:
        res, spp, views = 64, 4, 8
        scenes = [SceneBuild.workload(5, seed=7, variant=v, res=res) for v in range(views)]
        mats = [3, 4, 5, 6]
        per = []
        for v, sb in enumerate(scenes):
            ctx.upload(sb, 1 + v)
            per.append(DiffScene(ctx, sb, mats, slot=1 + v))
        ctx.upload(scenes[0], 0)
        base = DiffScene(ctx, scenes[0], mats, slot=0)
        multi = DiffViews(base, scenes)
        theta, em = base.get_params()
        rng = np.random.default_rng(3)
        theta2 = theta * rng.uniform(0.8, 1.2, theta.shape).astype(np.float32)
        em2 = em * rng.uniform(0.7, 1.3, em.shape).astype(np.float32)
        for d in [base] + per:
            d.set_params(theta2, em2, refresh_sampling=True)
        imgs = multi.render(40, spp)
        singles = np.stack([d.render(40, spp) for d in per])
        rel = np.abs(imgs - singles) / np.maximum(np.abs(singles), 1e-6)
        print(f"  images: max relative difference {rel.max():.2e}")
        assert imgs.shape == (views, res, res, 3)
        assert rel.max() < 1e-9
        adj = rng.standard_normal(imgs.shape).astype(np.float32)
        gs, ge = multi.adjoint(40, spp, adj)                 # reuses the forward's per-sample radiance
        gs_sum = np.zeros_like(gs)
        ge_sum = np.zeros_like(ge)
        for v, d in enumerate(per):
            g1, e1 = d.adjoint(40, spp, adj[v])
            gs_sum += g1
            ge_sum += e1
        es = np.abs(gs - gs_sum) / np.maximum(np.abs(gs_sum), 1e-12)
        ee = np.abs(ge - ge_sum).max() / np.abs(ge_sum).max()
        print(f"  gradients: scalars max relative difference {es.max():.2e}, EM max difference / max {ee:.2e}")
        assert np.abs(gs_sum).max() > 0 and np.abs(ge_sum).max() > 0
        assert es.max() < 1e-6 and ee < 1e-6
        # an adjoint of other iterations replays its own primal pass first: same gradients as the
        # per-slot adjoints of those iterations
        gs3, _ = multi.adjoint(80, spp, adj)
        gs3_sum = sum(d.adjoint(80, spp, adj[v])[0] for v, d in enumerate(per))
        assert np.max(np.abs(gs3 - gs3_sum) / np.maximum(np.abs(gs3_sum), 1e-12)) < 1e-6
    finally:
        ctx.close()
