"""Batched multi-scene renderer (pg3g_batch_*) against per-scene renders of the same iterations.

A batched pass traces every (pixel, iteration) sample exactly as pg3g_render_slot does, so the
per-pixel FP64 sums agree up to the order of the FP64 additions (work items group iterations
differently from the per-scene kernel's iteration chunks) and, rarely, one-ulp differences of
the separately compiled kernels that send a path another way.
"""
import numpy as np
import pytest

from pg3render_amd import Batch, Context, SceneBuild, _ffi
from pg3render_amd._ffi import Pg3gError

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def c3_scenes(n, res_lo=24, res_hi=72):
    """Small versions of the 64-scene batch workload (same materials, EMs, spp choices)."""
    rng = np.random.default_rng(5)
    out = []
    for v in range(n):
        sb = SceneBuild.workload(3, seed=2024, variant=v, res=int(rng.integers(res_lo, res_hi)))
        out.append(sb)
    return out


def plain_images(ctx, scenes, base_slot, passes):
    imgs = []
    for i, sb in enumerate(scenes):
        ctx.upload(sb, base_slot + i)
        ctx.render(0, sb.spp * passes, base_slot + i)
        imgs.append(ctx.framebuffer(base_slot + i, sb.resolution))
    return imgs


def assert_same(a, b, what):
    (ia, na), (ib, nb) = a, b
    assert na == nb, what
    rel = np.abs(ia - ib) / np.maximum(np.abs(ia), 1e-6)
    assert np.mean(rel > 1e-6) < 2e-3, (what, float(np.mean(rel > 1e-6)))
    assert abs(ia.sum() - ib.sum()) / max(abs(ia.sum()), 1e-12) < 1e-5, what


def test_batch_equals_per_scene_renders(ctx):
    scenes = c3_scenes(6)
    n = len(scenes)
    for i, sb in enumerate(scenes):
        ctx.upload(sb, i)
    batch = Batch(ctx, range(n), [sb.spp for sb in scenes])
    batch.set_config(iters_per_item=4, waves_per_cu=16, use_graph=True)
    batch.render(2)
    got = [ctx.framebuffer(i, sb.resolution) for i, sb in enumerate(scenes)]
    batch.close()
    want = plain_images(ctx, scenes, 100, passes=2)
    for i in range(n):
        assert_same(got[i], want[i], f"scene {i}")


def full_batch_scenes():
    """The 64-scene batch workload at its own sizes (sides 64..512, spp 4/16/64/256)."""
    return [SceneBuild.workload(3, seed=2024, variant=v) for v in range(64)]


def oracle_band_parity(sb, img, rows=8):
    """The batched image of one scene against the CPU oracle on a band of `rows` image rows at the
    scene's full spp (RunIteration x spp, path_tracer_base.hxx:57-90): the same RNG streams, so
    pixels agree to FP32 rounding except where rounding sends a path another way."""
    from oracle import oracle as O
    w, h = sb.resolution
    r0 = h // 2 - rows // 2
    ref, _ = O.OracleScene(sb).render(0, sb.spp, threads=0, pixel_begin=r0 * w, pixel_end=(r0 + rows) * w)
    g, o = img[r0:r0 + rows], ref[r0:r0 + rows]
    lum_g, lum_o = g.sum(-1), o.sum(-1)
    diff = np.abs(lum_g - lum_o) / np.maximum(np.abs(lum_o), 1e-3 * max(1e-6, lum_o.mean()))
    frac = float(np.mean(diff > 1e-3))
    band_rel = abs(lum_g.sum() - lum_o.sum()) / lum_o.sum()
    print(f"  {w}x{h} x {sb.spp} spp, rows {r0}..{r0 + rows - 1}: pixels off by >1e-3 {frac:.4f}, "
          f"max {diff.max():.2e}, band sum rel {band_rel:.2e}")
    assert frac < 0.03, frac
    assert band_rel < 5e-3, band_rel


def test_full_64_scene_batch_equals_per_scene_and_oracle(ctx):
    """The whole 64-scene workload (784.5 M samples) in one batched pass: every scene equals its
    per-scene render (pg3g_render_slot of the same iterations), and four scenes -- the largest
    (507^2 x 256 spp), the smallest, and one each of spp 4 and 64 -- equal the CPU oracle on a band
    of rows at full spp."""
    scenes = full_batch_scenes()
    slots = list(range(1, 65))
    for s, sb in zip(slots, scenes):
        ctx.upload(sb, s)
    batch = Batch(ctx, slots, [sb.spp for sb in scenes])
    batch.set_config(iters_per_item=8, waves_per_cu=16, use_graph=True)
    batch.render(1)
    got = [ctx.framebuffer(s, sb.resolution) for s, sb in zip(slots, scenes)]
    batch.close()
    want = plain_images(ctx, scenes, 100, passes=1)
    for i in range(len(scenes)):
        assert_same(got[i], want[i], f"scene {i}")
    work = [sb.resolution[0] * sb.resolution[1] * sb.spp for sb in scenes]
    picks = [int(np.argmax(work)), int(np.argmin(work))]
    for spp in (4, 64):
        cand = [i for i, sb in enumerate(scenes) if sb.spp == spp and i not in picks]
        picks.append(max(cand, key=lambda i: work[i]))
    assert scenes[picks[0]].resolution == (507, 507) and scenes[picks[0]].spp == 256
    for i in picks:
        img, it = got[i]
        assert it == scenes[i].spp
        print(f"scene {i}:")
        oracle_band_parity(scenes[i], img)


@pytest.mark.parametrize("waves", [4, 8, 32])
def test_batch_launch_shapes_equal_per_scene_renders(ctx, waves):
    """Every launch shape the autotuner can pick: workgroups of 256 and 512 lanes (4 and 8
    wavefronts per CU, one workgroup per CU) and two 1024-lane workgroups per CU (32, the second
    starting when the first has drained the work), with the per-XCD item parts, the stealing of
    items from other parts and the LDS cache of the scenes' steerable tables: the same images as
    per-scene renders."""
    scenes = c3_scenes(5)
    for i, sb in enumerate(scenes):
        ctx.upload(sb, i)
    batch = Batch(ctx, range(len(scenes)), [sb.spp for sb in scenes])
    batch.set_config(iters_per_item=2, waves_per_cu=waves, use_graph=True)
    batch.render(2)
    got = [ctx.framebuffer(i, sb.resolution) for i, sb in enumerate(scenes)]
    batch.close()
    want = plain_images(ctx, scenes, 100, passes=2)
    for i in range(len(scenes)):
        assert_same(got[i], want[i], f"scene {i}")


def test_batch_graph_and_direct_launch_agree(ctx):
    scenes = c3_scenes(3)
    for i, sb in enumerate(scenes):
        ctx.upload(sb, i)
    out = []
    for graph in (True, False):
        for i in range(len(scenes)):
            ctx.clear(i)
        b = Batch(ctx, range(len(scenes)), [sb.spp for sb in scenes])
        b.set_config(iters_per_item=8, waves_per_cu=8, use_graph=graph)
        b.render(1)
        out.append([ctx.framebuffer(i, sb.resolution) for i, sb in enumerate(scenes)])
        b.close()
    for i in range(len(scenes)):
        assert_same(out[0][i], out[1][i], f"scene {i}")


def test_autotune_restarts_and_keeps_results(ctx):
    scenes = c3_scenes(4)
    for i, sb in enumerate(scenes):
        ctx.upload(sb, i)
    b = Batch(ctx, range(len(scenes)), [sb.spp for sb in scenes])
    best = b.autotune(max_candidates=3)
    assert best > 0
    cfg = b.config()
    assert cfg["iters_per_item"] in (1, 2, 4, 8, 16) and cfg["waves_per_cu"] in (8, 16, 32)
    for i, sb in enumerate(scenes):
        img, it = ctx.framebuffer(i, sb.resolution)
        assert it == 0 and not img.any()
    b.render(1)
    got = [ctx.framebuffer(i, sb.resolution) for i, sb in enumerate(scenes)]
    b.close()
    want = plain_images(ctx, scenes, 100, passes=1)
    for i in range(len(scenes)):
        assert_same(got[i], want[i], f"scene {i}")


def test_mixed_light_setups_use_the_generic_kernel(ctx):
    """Scenes with different light setups (area lights only / EM + area lights / EM only)."""
    scenes = [SceneBuild.workload(1, seed=3, res=40), SceneBuild.workload(3, seed=2024, variant=7, res=36),
              SceneBuild.workload(2, seed=4, res=32)]
    spp = [3, 5, 2]
    for i, sb in enumerate(scenes):
        ctx.upload(sb, i)
    b = Batch(ctx, range(3), spp)
    b.set_config(iters_per_item=2, waves_per_cu=16)
    b.render(1)
    got = [ctx.framebuffer(i, sb.resolution) for i, sb in enumerate(scenes)]
    b.close()
    for i, sb in enumerate(scenes):
        ctx.upload(sb, 100 + i)
        ctx.render(0, spp[i], 100 + i)
        assert_same(got[i], ctx.framebuffer(100 + i, sb.resolution), f"scene {i}")


def test_batch_is_invalidated_by_uploads_and_rejects_bad_input(ctx):
    sb = SceneBuild.workload(3, seed=2024, variant=1, res=24)
    ctx.upload(sb, 0)
    with pytest.raises(Pg3gError):
        Batch(ctx, [0, 0], [4, 4])            # duplicate slot
    with pytest.raises(Pg3gError):
        Batch(ctx, [0], [0])                  # zero spp
    with pytest.raises(Pg3gError):
        Batch(ctx, [250], [4])                # empty slot
    b = Batch(ctx, [0], [4])
    with pytest.raises(Pg3gError):
        b.set_config(iters_per_item=0)
    ctx.upload(sb, 1)
    with pytest.raises(Pg3gError) as e:
        b.render(1)
    assert e.value.code == _ffi.PG3G_ERR_NO_SCENE
    b.close()
    ptn = SceneBuild.workload(1, seed=3, res=16)
    ptn.set_params(algorithm="ptn")
    ctx.upload(ptn, 2)
    with pytest.raises(Pg3gError) as e:
        Batch(ctx, [2], [1])
    assert e.value.code == _ffi.PG3G_ERR_UNSUPPORTED


def test_batch_pass_and_parameter_change_launch_queued_run_iterations(ctx):
    """Iterations queued by pg3g_run_iteration (RunIteration, renderer.hxx:232-278) are launched
    before anything else acts on their slot: a batched pass of the slot adds to them, and a
    parameter change leaves them rendered with the parameters they were queued under."""
    scenes = c3_scenes(2)
    for i, sb in enumerate(scenes):
        ctx.upload(sb, i)
    for i in range(len(scenes)):
        for it in range(3):
            ctx.run_iteration(it, i)                          # queued, not launched
    b = Batch(ctx, range(len(scenes)), [sb.spp for sb in scenes])
    b.set_config(iters_per_item=2, waves_per_cu=16, use_graph=True)
    b.render(1)                                               # pass 0: iterations [0, spp)
    got = [ctx.framebuffer(i, sb.resolution) for i, sb in enumerate(scenes)]
    b.close()
    for i, sb in enumerate(scenes):
        ctx.upload(sb, 100 + i)
        ctx.render(0, 3, 100 + i)
        ctx.render(0, sb.spp, 100 + i)
        assert got[i][1] == 3 + sb.spp
        assert_same(got[i], ctx.framebuffer(100 + i, sb.resolution), f"scene {i}")

    sb = scenes[0]
    ctx.upload(sb, 0)
    for it in range(5):
        ctx.run_iteration(it, 0)
    changed = type(sb.desc.params).from_buffer_copy(sb.desc.params)
    changed.max_path_length = 2
    ctx.set_params(0, changed)                                # launches iterations 0..4 first
    img, n = ctx.framebuffer(0, sb.resolution)
    ctx.upload(sb, 100)
    ctx.render(0, 5, 100)
    assert n == 5
    assert_same((img, n), ctx.framebuffer(100, sb.resolution), "parameters of the queued iterations")
