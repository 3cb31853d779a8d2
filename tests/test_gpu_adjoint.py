"""Path-replay adjoint (pg3g_adjoint) against central finite differences.

Scene: the differentiation workload (box, two GGX-conductor spheres with RGB tint and roughness,
128x64 sky EM sampled by the simple spherical sampler) at a reduced resolution.  Loss: mean
squared error against a target rendered from hidden parameters.  Finite differences use the
same samples (correlated RNG: identical iteration range) and frozen sampling, with the loss
accumulated in FP64 on the CPU, h = 1e-3.  64 parameters: all 11 scalars (two conductor tints and
roughnesses, the white wall's albedo) and 53 EM texels drawn at random among those the image
depends on.  Required agreement: 2 %.
"""
import numpy as np
import pytest

from pg3render_amd import Context, SceneBuild
from pg3render_amd.diff import DiffScene, render_function

pytestmark = pytest.mark.gpu

RES, SPP, H = 48, 8, 1e-3
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


def test_adjoint_matches_central_differences(setup):
    _, _, ds, theta0, em0, target = setup
    img = ds.render(0, SPP)
    dl = 2.0 * (img - target) / img.size
    gs, ge = ds.adjoint(0, SPP, dl)
    rng = np.random.default_rng(7)
    live = np.flatnonzero(np.abs(ge).ravel() > 1e-3 * np.abs(ge).max())
    texels = rng.choice(live, 53, replace=False)
    params = [("s", i) for i in range(11)] + [("e", int(t)) for t in texels]
    rows = []
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
        fd = (ls[0] - ls[1]) / (2 * H)
        g = gs[i] if kind == "s" else ge.reshape(-1)[i]
        rows.append((kind, i, g, fd))
    ds.set_params(theta0, em0, refresh_sampling=True)
    g = np.array([r[2] for r in rows])
    fd = np.array([r[3] for r in rows])
    scale = np.maximum(np.abs(fd), np.abs(g))
    floor = 1e-3 * scale.max()
    rel = np.abs(g - fd) / np.maximum(scale, floor)
    worst = int(np.argmax(rel))
    print(f"64 parameters: median relative error {np.median(rel):.2e}, max {rel.max():.2e} at {rows[worst]}")
    assert rel.max() <= 0.02, rows[worst]


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
