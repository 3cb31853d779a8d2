"""GPU path (libpg3g.so through the C ABI) against the CPU oracle on identical seeded inputs.

Function level: BSDF eval/pdf/sample and EM eval/sample/pdf agree to <= 1e-5 relative (FP32,
north_star tier 1).  Image level: with the same counter-based RNG both sides trace the same paths,
so per-pixel sums agree except where FP32 rounding sends a path another way.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from pg3render_amd import Context, GpuPathTracer, SceneBuild, _ffi
from pg3render_amd.scene_desc import CustomScene, light, material

pytestmark = pytest.mark.gpu

REL = 1e-5      # tier-1 tolerance (relative, FP32)


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def tier1_report(what, e):
    """Prints the maximum relative error and how many samples exceed the 1e-5 bound."""
    e = np.asarray(e, np.float64).ravel()
    print(f"{what}: max {e.max():.2e}, q999 {np.quantile(e, 0.999):.2e}, above 1e-5: {int(np.sum(e > REL))} of {e.size}")


def rel_err(a, b, floor):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), floor)


def dirs(rng, n, both=False):
    u = rng.random((n, 2))
    z = u[:, 1] * 0.998 + 0.001
    r = np.sqrt(1 - z * z)
    d = np.stack([r * np.cos(2 * np.pi * u[:, 0]), r * np.sin(2 * np.pi * u[:, 0]), z], 1)
    if both:
        d[rng.random(n) < 0.5, 2] *= -1
    return d.astype(np.float32)


MATERIALS = [
    ("lambert", material(_ffi.MAT_LAMBERT, diffuse=(0.6, 0.3, 0.9)), False),
    ("phong", material(_ffi.MAT_PHONG, diffuse=(0.8, 0.8, 0.15), glossy=(0.7, 0.7, 0.7), phong_exponent=200,
                       phong_diffuse_on=1, phong_glossy_on=1), False),
    ("ggx_conductor_0.1", material(_ffi.MAT_GGX_CONDUCTOR, alpha=0.1, inner_ior=0.2, absorbance=3.8157), False),
    ("ggx_conductor_0.5", material(_ffi.MAT_GGX_CONDUCTOR, alpha=0.5, inner_ior=0.468, outer_ior=1.000277,
                                   absorbance=2.81), False),
    ("ggx_dielectric_0.3", material(_ffi.MAT_GGX_DIELECTRIC, alpha=0.3, inner_ior=1.5), True),
    ("ggx_dielectric_0.01", material(_ffi.MAT_GGX_DIELECTRIC, alpha=0.01, inner_ior=1.51, outer_ior=1.000277), True),
    ("smooth_conductor", material(_ffi.MAT_SMOOTH_CONDUCTOR, inner_ior=0.121, absorbance=3.66), False),
    ("smooth_dielectric", material(_ffi.MAT_SMOOTH_DIELECTRIC, inner_ior=1.51, outer_ior=1.000277), True),
]


@pytest.mark.parametrize("name,mat,both", MATERIALS, ids=[m[0] for m in MATERIALS])
def test_bsdf_eval_parity(ctx, name, mat, both):
    rng = np.random.default_rng(hash(name) % 2**32)
    n = 20000
    wil, wol = dirs(rng, n, both), dirs(rng, n, both)
    fg, pg, cg = ctx.bsdf_eval(mat, wil, wol)
    fo, po, co = O.bsdf_eval(mat, wil, wol)
    scale_f = max(1e-3, float(np.abs(fo).max()) * 1e-6)
    ef = rel_err(fg, fo, scale_f)
    ep = rel_err(pg, po, max(1e-3, float(np.abs(po).max()) * 1e-6))
    assert np.all(cg == co)
    tier1_report(f"{name} eval f", ef)
    tier1_report(f"{name} eval pdf", ep)
    # every sample within the tier-1 bound (measured: at most 3e-7)
    assert ef.max() <= REL and ep.max() <= REL, (ef.max(), ep.max())


@pytest.mark.parametrize("name,mat,both", MATERIALS, ids=[m[0] for m in MATERIALS])
def test_bsdf_sample_parity(ctx, name, mat, both):
    """SampleBsdf = warp (uniforms -> direction, component) + value (attenuation, pdf).  The warp is
    compared directly; the value is compared with the oracle's EvalBsdf at the direction the GPU
    sampled, which is what SampleBsdf returns in the reference for every finite BSDF (the GGX
    materials use the sampled microfacet normal instead of the recomputed halfway vector, equal up
    to rounding).  Comparing values at two different directions would measure the conditioning of
    the BSDF (e.g. cos^200 for the Phong lobe), not the implementation."""
    rng = np.random.default_rng(1 + hash(name) % 2**32)
    n = 20000
    wol = dirs(rng, n, both)
    u = rng.random((n, 3)).astype(np.float32)
    wg, fg, pg, cg = ctx.bsdf_sample(mat, wol, u)
    wo_, fo, po, co = O.bsdf_sample(mat, wol, u)
    finite = np.isfinite(po)
    assert np.array_equal(finite, np.isfinite(pg))
    # warp: unit directions.  A one-ulp input difference legitimately switches a branch at the
    # reference's own discontinuities (Frame::SetFromZ changes its helper axis at |z.x| = 0.99,
    # math.hxx:983-989; total internal reflection; the horizon), so those samples are counted
    # separately and must stay rare.
    ew = np.abs(wg - wo_).max(axis=1)
    flips = ew > 1e-3
    q_warp = float(np.quantile(ew[~flips], 0.999))
    stats = f"{name}: flips {flips.mean():.2e}, warp q999 {q_warp:.2e}"
    tier1_report(f"{name} sample direction (no branch flip)", ew[~flips])
    assert flips.mean() < 5e-3, stats
    assert q_warp <= REL, stats
    fe, pe, ce = O.bsdf_eval(mat, wg, wol)
    live = finite & (fg.max(axis=1) > 0) & (fe.max(axis=1) > 0) & ~flips
    if live.any():
        # value at the GPU-sampled direction
        ef = rel_err(fg[live], fe[live], 1e-30).max(axis=1)
        ep = rel_err(pg[live], pe[live], 1e-30)
        tier1_report(f"{name} sampled value at the sampled direction", ef)
        tier1_report(f"{name} sampled pdf at the sampled direction", ep)
        stats += f", value-at-direction q999 f {np.quantile(ef, 0.999):.2e} pdf {np.quantile(ep, 0.999):.2e}"
        if mat.type in (_ffi.MAT_LAMBERT, _ffi.MAT_PHONG):
            # SampleBsdf returns EvalBsdf at the sampled direction (materials.hxx:285-292, 494-503)
            assert np.quantile(ef, 0.999) <= REL and np.quantile(ep, 0.999) <= REL, stats
        else:
            # GGX SampleBsdf evaluates at the sampled microfacet normal, EvalBsdf at the halfway
            # vector recomputed from the two directions; the reference's own two results differ by
            # the conditioning of D(h) (~1/alpha).  Direct sample-to-sample comparison where both
            # warps produced the same FP32 direction is held to the tier-1 bound.
            same = live & (ew <= 1.2e-7)
            if same.sum() > 100:
                es = np.maximum(rel_err(fg[same], fo[same], 1e-30).max(axis=1), rel_err(pg[same], po[same], 1e-30))
                stats += f", same-direction q99 {np.quantile(es, 0.99):.2e} ({same.sum()} samples)"
                assert np.quantile(es, 0.99) <= REL, stats
            # D(h) at alpha 0.01 amplifies the ~1e-7 halfway rounding by ~(1/alpha)^2
            tol = 1e-3 if mat.alpha >= 0.1 else 1e-3 * (0.1 / mat.alpha) ** 2
            assert np.quantile(np.maximum(ef, ep), 0.99) <= tol, stats
    print(stats)
    # Dirac components: values compared directly (no direction dependence beyond the branch)
    if (~finite).any():
        d = ~finite & ~flips
        assert np.quantile(rel_err(fg[d], fo[d], 1e-6).max(axis=1), 0.999) <= REL
        assert np.quantile(rel_err(cg[d], co[d], 1e-6), 0.999) <= REL


def em_scene(sampler, w=256, h=128, seed=3, levels=(4, 4)):
    rng = np.random.default_rng(seed)
    yy, xx = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    img = 0.2 + 0.8 * (1 - yy / h)[..., None] * np.array([0.5, 0.75, 1.0])
    img = img + 200.0 * np.exp(-((xx - w * 0.3) ** 2 + (yy - h * 0.3) ** 2) / 8.0)[..., None]
    img = img * (0.9 + 0.2 * rng.random((h, w, 3)))
    return CustomScene(materials=[material(_ffi.MAT_LAMBERT, diffuse=(0.5, 0.5, 0.5))],
                       lights=[light(_ffi.LIGHT_BACKGROUND)], em=img.astype(np.float32), em_sampler=sampler,
                       steerable_levels=levels, camera={"res": (8, 8)})


@pytest.mark.parametrize("sampler", [_ffi.EM_SIMPLE, _ffi.EM_STEERABLE, _ffi.EM_COSINE])
def test_em_parity(ctx, sampler):
    sc = em_scene(sampler)
    ctx.upload(sc)
    orc = O.OracleScene(sc)
    rng = np.random.default_rng(17)
    n = 20000
    normals = dirs(rng, n)
    normals = normals[:, [0, 2, 1]] * np.where(rng.random((n, 1)) < 0.5, -1, 1).astype(np.float32)
    sides = rng.integers(1, 4, n).astype(np.int32)
    u = rng.random((n, 3)).astype(np.float32)
    # evaluation
    d = dirs(rng, n, True)
    eg, eo = ctx.em_eval(d), orc.em_eval(d)
    e_eval = rel_err(eg, eo, 1e-3).max(axis=1)
    tier1_report(f"sampler {sampler} EM eval", e_eval)
    # Every sample: within 1e-5 plus what the oracle's own value moves when its lat-long
    # coordinates move by up to 2 ulp (the GPU's atan2/acos and the host's differ by an ulp or so;
    # next to the sun's steep gradient that ulp moves the tent-filtered value by more than 1e-5).
    uv = np.array([O.dir2latlong(x) for x in d], np.float32)
    env = np.zeros(n)
    for axis in (0, 1):
        for k in (-2, -1, 1, 2):
            q = uv.copy()
            q[:, axis] = np.nextafter(q[:, axis], np.float32(np.sign(k) * 2.0))
            if abs(k) == 2:
                q[:, axis] = np.nextafter(q[:, axis], np.float32(np.sign(k) * 2.0))
            moved = rel_err(orc.em_eval_uv(np.clip(q, 0.0, 1.0)), eo, 1e-3).max(axis=1)
            env = np.maximum(env, moved)
    excess = e_eval - (REL + env)
    print(f"sampler {sampler} EM eval: worst excess over 1e-5 + 2-ulp envelope {excess.max():.2e}, "
          f"samples above 1e-5 {int(np.sum(e_eval > REL))}, all within the envelope")
    assert np.all(excess <= 0.0), (e_eval[excess > 0], env[excess > 0])
    # sampling
    dg, pg, rg, okg = ctx.em_sample(normals, sides, u)
    do, po, ro, oko = orc.em_sample(normals, sides, u)
    assert np.all(okg == oko)
    good = oko == 1
    ed = np.abs(dg - do).max(axis=1)[good]
    ep = rel_err(pg, po, 1e-6)[good]
    er = rel_err(rg, ro, 1e-3).max(axis=1)[good]
    for nm, e in (("direction", ed), ("pdf", ep), ("radiance*cos", er)):
        tier1_report(f"sampler {sampler} EM sample {nm}", e)
    # sampled directions: every sample (measured: at most 5e-7); pdf and radiance*cos inherit the
    # one-ulp differences of the lat-long mapping's atan2/acos next to the sun's steep gradient
    assert ed.max() <= REL, ed.max()
    for e in (ep, er):
        assert np.quantile(e, 0.995) <= REL * 2, e.max()
        assert np.mean(e > 1e-3) < 2e-3
    # pdf of arbitrary directions.  The simple sampler's pdf is a difference of adjacent FP32 CDF
    # entries times the row width; the GPU's parallel FP64 scan adds in another order than the
    # oracle's sequential sum, so an entry occasionally rounds to the neighbouring float, and one
    # ulp (6e-8) over an entry difference of ~1/256 is ~1.5e-5 of the pdf.  Printed: the tail.
    pg2 = ctx.em_pdf(d, normals, sides)
    po2 = orc.em_pdf(d, normals, sides)
    e = rel_err(pg2, po2, 1e-6)
    tier1_report(f"sampler {sampler} EM pdf", e)
    assert np.quantile(e, 0.999) <= REL * 2 and np.mean(e > 1e-3) < 1e-3 and e.max() < 1e-4


def test_steerable_tables_match_oracle_build(ctx):
    sb = SceneBuild.workload(2, seed=5)
    ctx.upload(sb)
    leaves_g, nv_g, levels, wg = ctx.steerable_leaves()
    leaves_o, nv_o, wo = O.OracleScene(sb).steerable()
    assert leaves_g.shape == leaves_o.shape == (5120, 3, 3)
    assert nv_g == nv_o
    assert np.max(np.abs(leaves_g - leaves_o)) <= 1e-6
    assert levels == 6   # ceil(log6 5120) + 1 (em_steerable_sampler.hxx:4518-4628)
    assert np.allclose(wg, wo, rtol=1e-5, atol=1e-5 * np.abs(wo).max())


def render_pair(ctx, sb, spp, first=0):
    ctx.upload(sb)
    ctx.render(first, spp)
    g, it = ctx.framebuffer(0, sb.resolution)
    o, _ = O.OracleScene(sb).render(first, spp, threads=8)
    assert it == spp
    return g, o


def assert_image_parity(g, o, frac=0.02):
    assert np.isfinite(g).all()
    lum_g, lum_o = g.sum(-1), o.sum(-1)
    diff = np.abs(lum_g - lum_o) / np.maximum(np.abs(lum_o), 1e-3 * max(1e-6, lum_o.mean()))
    # identical RNG streams: most pixels agree to FP32 rounding
    assert np.mean(diff > 1e-3) < frac, np.mean(diff > 1e-3)
    assert abs(lum_g.sum() - lum_o.sum()) / lum_o.sum() < 5e-3


KERNELS = [pytest.param(_ffi.KERNEL_MEGAKERNEL, id="megakernel"), pytest.param(_ffi.KERNEL_WAVEFRONT, id="wavefront")]


@pytest.mark.parametrize("kernel", KERNELS)
def test_workload1_image_parity(ctx, kernel):
    ctx.set_kernel_mode(kernel)
    sb = SceneBuild.workload(1, seed=21, res=48)
    g, o = render_pair(ctx, sb, 8)
    assert_image_parity(g, o)


@pytest.mark.parametrize("kernel", KERNELS)
def test_workload2_image_parity(ctx, kernel):
    ctx.set_kernel_mode(kernel)
    sb = SceneBuild.workload(2, seed=22, res=40)
    g, o = render_pair(ctx, sb, 4)
    assert_image_parity(g, o, frac=0.03)


FULL_SIZE = [pytest.param(1, {}, _ffi.KERNEL_MEGAKERNEL, id="C1_box_area_light_256px_16spp"),
             pytest.param(2, {}, _ffi.KERNEL_MEGAKERNEL, id="C2_box_steerable_em_512px_64spp"),
             pytest.param(2, {}, _ffi.KERNEL_WAVEFRONT, id="C2_box_steerable_em_512px_64spp_wavefront"),
             pytest.param(2, dict(splitting_budget=4.0, indirect_clipping=16.0, max_path_length=0),
                          _ffi.KERNEL_MEGAKERNEL, id="C2_reference_defaults_sb4_iic16_rr")]


@pytest.mark.parametrize("workload,params,kernel", FULL_SIZE)
def test_full_size_frame_against_the_oracle(ctx, workload, params, kernel):
    """The bench's workloads at their full sizes (C1 256x256 x 16 spp, C2 512x512 x 64 spp through
    the fused kernel and the wavefront pipeline, and C2 at the reference's defaults): the whole
    frame on the GPU, and a band of 8 image rows through the oracle at the same spp (RunIteration x
    spp, path_tracer_base.hxx:57-90; the same RNG streams, so pixels agree to FP32 rounding except
    where rounding sends a path another way).  The whole frame is finite."""
    ctx.set_kernel_mode(kernel)
    sb = SceneBuild.workload(workload, seed=2024)
    sb.set_params(**params)
    ctx.upload(sb)
    ctx.render(0, sb.spp)
    img, it = ctx.framebuffer(0, sb.resolution)
    w, h = sb.resolution
    assert it == sb.spp and img.shape == (h, w, 3)
    assert np.isfinite(img).all() and img.sum() > 0
    r0 = h // 2 - 4
    ref, _ = O.OracleScene(sb).render(0, sb.spp, threads=0, pixel_begin=r0 * w, pixel_end=(r0 + 8) * w)
    g, o = img[r0:r0 + 8], ref[r0:r0 + 8]
    lum_g, lum_o = g.sum(-1), o.sum(-1)
    diff = np.abs(lum_g - lum_o) / np.maximum(np.abs(lum_o), 1e-3 * max(1e-6, lum_o.mean()))
    print(f"{w}x{h} x {sb.spp} spp, rows {r0}..{r0 + 7}: pixels off by >1e-3 {np.mean(diff > 1e-3):.4f}, "
          f"max {diff.max():.2e}, band sum rel {abs(lum_g.sum() - lum_o.sum()) / lum_o.sum():.2e}")
    assert_image_parity(g, o, frac=0.03)


@pytest.mark.parametrize("workload,sampler", [(2, None), (2, _ffi.EM_SIMPLE), (2, _ffi.EM_COSINE), (1, None), (3, None)],
                         ids=["steerable_em", "simple_em", "cosine_em", "area_lights", "lights_and_em"])
def test_wavefront_equals_megakernel(ctx, workload, sampler):
    """Both kernel organisations run the same per-vertex code on the same RNG streams: the
    one-light steerable vertex, the simple sampler with its LDS-staged marginal, the cosine
    sampler, area lights only, and area lights with an EM (light picking)."""
    sb = SceneBuild.workload(workload, seed=25, res=64)
    if sampler is not None:
        sb.desc.em.sampler = sampler
    out = []
    for k in (_ffi.KERNEL_MEGAKERNEL, _ffi.KERNEL_WAVEFRONT):
        ctx.set_kernel_mode(k)
        ctx.upload(sb)
        ctx.render(0, 8)
        out.append(ctx.framebuffer(0, sb.resolution)[0])
    a, b = out
    rel = np.abs(a - b) / np.maximum(np.abs(a), 1e-6)
    assert np.mean(rel > 1e-6) < 1e-3
    assert abs(a.sum() - b.sum()) / a.sum() < 1e-6


@pytest.mark.parametrize("build", [pytest.param(lambda: SceneBuild.workload(2, seed=28, res=48), id="workload2"),
                                   pytest.param(lambda: SceneBuild.reference(33, resolution=32), id="reference_scene33"),
                                   pytest.param(lambda: SceneBuild.reference(12, resolution=32), id="reference_scene12")])
def test_wavefront_russian_roulette(ctx, build):
    """-maxpl 0: the wavefront pipeline runs bounces until no path is live (the host reads the
    live count back every 8 bounces) and matches the fused kernel on the same RNG streams, also
    where paths bounce inside a dielectric sphere for dozens of vertices (scene 33)."""
    sb = build()
    sb.set_params(algorithm="pt", splitting_budget=1.0, indirect_clipping=0.0, max_path_length=0, seed=5)
    out = []
    for k in (_ffi.KERNEL_MEGAKERNEL, _ffi.KERNEL_WAVEFRONT):
        ctx.set_kernel_mode(k)
        ctx.upload(sb)
        ctx.synchronize()
        ctx.reset_counters()
        ctx.render(0, 8)
        out.append((ctx.framebuffer(0, sb.resolution)[0], ctx.counters()))
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)
    (a, ca), (b, cb) = out
    assert ca == cb                      # the same segments and samples
    rel = np.abs(a - b) / np.maximum(np.abs(a), 1e-6)
    assert np.mean(rel > 1e-6) < 1e-3
    assert abs(a.sum() - b.sum()) / a.sum() < 1e-6


@pytest.mark.parametrize("res", [37, 9])
def test_wavefront_ragged_resolution(ctx, res):
    """Resolutions that leave 8x8 tiles partly outside the image: padding samples are never
    enqueued and every pixel is accumulated once per sample."""
    sb = SceneBuild.workload(2, seed=27, res=res)
    out = []
    for k in (_ffi.KERNEL_MEGAKERNEL, _ffi.KERNEL_WAVEFRONT):
        ctx.set_kernel_mode(k)
        ctx.upload(sb)
        ctx.synchronize()
        ctx.reset_counters()
        ctx.render(0, 6)
        out.append(ctx.framebuffer(0, sb.resolution)[0])
        assert ctx.counters()[1] == res * res * 6
    a, b = out
    assert a.shape[:2] == (res, res)
    rel = np.abs(a - b) / np.maximum(np.abs(a), 1e-6)
    assert np.mean(rel > 1e-6) < 1e-3
    assert abs(a.sum() - b.sum()) / a.sum() < 1e-6


def test_wavefront_batch_larger_than_queue_capacity(ctx):
    """A render of more samples than one wavefront batch holds (2^21): two batches, partial tail."""
    ctx.set_kernel_mode(_ffi.KERNEL_WAVEFRONT)
    sb = SceneBuild.workload(1, seed=26, res=200)
    g, o = render_pair(ctx, sb, 60)
    assert_image_parity(g, o)


@pytest.mark.parametrize("alg", ["ptn", "dbs", "dlss", "dlsa", "dmis"])
def test_algorithms_image_parity(ctx, alg):
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)
    sb = SceneBuild.workload(1, seed=23, res=40)
    sb.set_params(algorithm=alg)
    g, o = render_pair(ctx, sb, 8)
    assert_image_parity(g, o)


@pytest.mark.parametrize("scene_id", [2, 3, 11, 18, 32, 34, 37])
def test_reference_scene_parity(ctx, scene_id):
    """The reference's own Cornell-box scene IDs with -sb 1 -iic 0 -maxpl 8."""
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)
    sb = SceneBuild.reference(scene_id, em_type=0, resolution=40)
    sb.set_params(splitting_budget=1.0, indirect_clipping=0.0, max_path_length=8, seed=7)
    g, o = render_pair(ctx, sb, 4)
    assert_image_parity(g, o, frac=0.03)


@pytest.mark.parametrize("scene_id,em_type", [(22, 2), (22, 3), (22, 7), (12, 3)])
def test_reference_scene_image_em_parity(ctx, scene_id, em_type, tmp_path):
    """Reference scenes 22 (environment light only) and 12 (four lights) with image EMs: the
    constant white 8x4 and 1024x512 images built in memory (em types 2, 3), and a 64x32 probe
    read from a PFM file with the reference's rotation 0.83 and scale 1.5 of em type 7
    (scene.hxx:891-897); simple spherical sampler, the reference's default."""
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)
    path = None
    if em_type == 7:
        rng = np.random.default_rng(9)
        img = (rng.random((32, 64, 3)) ** 4 * 3).astype(np.float32)
        path = str(tmp_path / "probe.pfm")
        with open(path, "wb") as f:
            f.write(b"PF\n64 32\n-1.0\n" + np.ascontiguousarray(img[::-1]).astype("<f4").tobytes())
    sb = SceneBuild.reference(scene_id, em_type=em_type, resolution=40, em_path=path)
    assert sb.desc.em.sampler == _ffi.EM_SIMPLE
    sb.set_params(splitting_budget=1.0, indirect_clipping=0.0, max_path_length=8, seed=7)
    g, o = render_pair(ctx, sb, 4)
    assert_image_parity(g, o, frac=0.03)


def test_iterations_accumulate_like_run_iteration(ctx):
    """Rendering iterations [0,4) then [4,8) equals rendering [0,8) (RunIteration is additive)."""
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)
    sb = SceneBuild.workload(1, seed=24, res=32)
    ctx.upload(sb)
    ctx.render(0, 4)
    ctx.render(4, 4)
    a, it = ctx.framebuffer(0, sb.resolution)
    ctx.clear(0)
    ctx.render(0, 8)
    b, _ = ctx.framebuffer(0, sb.resolution)
    assert it == 8
    assert np.allclose(a, b, rtol=1e-12, atol=1e-12)


def test_queued_run_iterations_equal_one_launch(ctx):
    """pg3g_run_iteration, RunIteration called once per iteration (renderer.hxx:232-278): 300
    consecutive calls (a run of PG3G_ITERATION_RUN = 256 launched by the 256th, the other 44 by
    the last call) equal one launch of iterations [0,300); a read launches what is still queued; an
    iteration that does not follow the queued ones launches them first; a clear drops what is
    queued."""
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)
    sb = SceneBuild.workload(1, seed=24, res=32)

    def plain(ranges):
        ctx.upload(sb, 7)
        for first, n in ranges:
            ctx.render(first, n, 7)
        return ctx.framebuffer(7, sb.resolution)

    ctx.upload(sb, 6)
    for i in range(300):
        ctx.run_iteration(i, 6, launch_now=(i == 299))
    a, it = ctx.framebuffer(6, sb.resolution)
    b, _ = plain([(0, 300)])
    assert it == 300
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)

    ctx.clear(6)
    for i in list(range(0, 5)) + list(range(10, 15)):      # a gap after iteration 4
        ctx.run_iteration(i, 6)
    a, it = ctx.framebuffer(6, sb.resolution)              # the read launches 10..14
    b, _ = plain([(0, 5), (10, 5)])
    assert it == 10
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)

    ctx.clear(6)
    for i in range(3):
        ctx.run_iteration(i, 6)
    ctx.clear(6)                                           # drops the three queued iterations
    a, it = ctx.framebuffer(6, sb.resolution)
    assert it == 0 and not a.any()

    tracer = GpuPathTracer(sb, context=ctx, slot=6)        # the AbstractRenderer-shaped loop
    for i in range(70):
        tracer.run_iteration(i)
    img = tracer.get_framebuffer()
    b, _ = plain([(0, 70)])
    np.testing.assert_allclose(img, b / 70, rtol=1e-12, atol=1e-12)


def test_unsupported_parameters_are_rejected(ctx):
    sb = SceneBuild.workload(1, seed=1, res=16)
    sb.set_params(splitting_budget=100.0)
    with pytest.raises(_ffi.Pg3gError) as e:
        ctx.upload(sb)
    assert e.value.code == _ffi.PG3G_ERR_UNSUPPORTED
    sb.set_params(splitting_budget=0.5)
    with pytest.raises(_ffi.Pg3gError) as e:
        ctx.upload(sb)
    assert e.value.code == _ffi.PG3G_ERR_UNSUPPORTED


# Splitting (-sb) and indirect clipping (-iic), path_tracer.hxx:167-454: the GPU walks the same
# recursion (explicit vertex stack) with the same per-branch RNG streams as the oracle.
SPLIT_CASES = [
    pytest.param(dict(splitting_budget=4.0, indirect_clipping=16.0, max_path_length=8), id="sb4_iic16_maxpl8"),
    pytest.param(dict(splitting_budget=4.0, indirect_clipping=0.0, max_path_length=0), id="sb4_rr"),
    pytest.param(dict(splitting_budget=1.0, indirect_clipping=0.5, max_path_length=8), id="sb1_iic0.5"),
    pytest.param(dict(splitting_budget=16.0, splitting_level=0.5, light_to_bsdf_ratio=2.0, indirect_clipping=4.0,
                      max_path_length=6), id="sb16_level0.5_ratio2"),
    pytest.param(dict(splitting_budget=4.0, indirect_clipping=16.0, max_path_length=0), id="defaults_sb4_iic16_rr"),
]


@pytest.mark.parametrize("params", SPLIT_CASES)
def test_splitting_and_clipping_parity(ctx, params):
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)
    sb = SceneBuild.workload(1, seed=31, res=32)
    sb.set_params(**params)
    g, o = render_pair(ctx, sb, 4)
    assert_image_parity(g, o, frac=0.03)


STEERABLE_SPLIT_CASES = [
    pytest.param(dict(splitting_budget=4.0, indirect_clipping=16.0, max_path_length=0), id="defaults_sb4_iic16_rr"),
    pytest.param(dict(splitting_budget=4.0, indirect_clipping=16.0, max_path_length=8), id="sb4_iic16_maxpl8"),
    # two light samples over four sub-samples: sub-samples 2 and 3 take none
    pytest.param(dict(splitting_budget=4.0, light_to_bsdf_ratio=0.5, indirect_clipping=4.0, max_path_length=0),
                 id="sb4_ratio0.5_rr"),
    # two light samples per sub-sample: the generic split kernel
    pytest.param(dict(splitting_budget=3.0, light_to_bsdf_ratio=2.0, indirect_clipping=16.0, max_path_length=0),
                 id="sb3_ratio2_rr"),
]


@pytest.mark.parametrize("params", STEERABLE_SPLIT_CASES)
def test_split_path_with_steerable_em_parity(ctx, params):
    """First-vertex splitting on the box with a steerable EM as its only light, the workload of the
    reference's defaults: the split kernel whose lanes share the EM lookups of escaped paths and
    light samples (one light sample per lane and vertex), and the generic split kernel for light
    sample counts above the BSDF sample count."""
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)
    sb = SceneBuild.workload(2, seed=23, res=32)
    sb.set_params(**params)
    g, o = render_pair(ctx, sb, 4)
    assert_image_parity(g, o, frac=0.03)


@pytest.mark.parametrize("scene_id", [33, 35, 36])
def test_tree_walker_keeps_the_1200_vertex_limit(ctx, scene_id):
    """Splitting below the first vertex (-sb 4 -sl 0.5: three BSDF samples at the first vertex,
    then chains) under Russian roulette on the reference scenes with a GGX dielectric sphere, where
    paths bounce inside the sphere for many vertices: the general tree walker against the oracle,
    which cuts paths at the reference's 1200 vertices (path_tracer.hxx:232-238)."""
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)
    sb = SceneBuild.reference(scene_id, resolution=24)
    sb.set_params(algorithm="pt", splitting_budget=4.0, splitting_level=0.5, indirect_clipping=16.0,
                  max_path_length=0, seed=13)
    g, o = render_pair(ctx, sb, 4)
    assert_image_parity(g, o, frac=0.03)


def closed_white_box(albedo=0.99, res=16):
    """A closed box of Lambert walls of albedo 0.99 (Russian roulette keeps a path with probability
    0.99 per vertex: chains of ~100 vertices, far beyond the 32 fold records a lane keeps) lit by a
    small one-sided area light under the ceiling; the camera is inside."""
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.float32)
    faces = [(0, 1, 2, 3), (4, 7, 6, 5), (0, 4, 5, 1), (1, 5, 6, 2), (2, 6, 7, 3), (3, 7, 4, 0)]   # inward
    tris = []
    for a, b, c, d in faces:
        tris += [np.concatenate([v[a], v[b], v[c]]), np.concatenate([v[a], v[c], v[d]])]
    # light quad just below the ceiling, facing down (-z)
    l0, l1, l2, l3 = (0.4, 0.4, 0.99), (0.6, 0.4, 0.99), (0.6, 0.6, 0.99), (0.4, 0.6, 0.99)
    tris += [np.array(l0 + l2 + l1, np.float32), np.array(l0 + l3 + l2, np.float32)]
    mats = [material(_ffi.MAT_LAMBERT, diffuse=(albedo,) * 3),
            material(_ffi.MAT_LAMBERT, diffuse=(0.0, 0.0, 0.0), light_id=0),
            material(_ffi.MAT_LAMBERT, diffuse=(0.0, 0.0, 0.0), light_id=1)]
    lights = [light(_ffi.LIGHT_AREA, p0=l0, p1=l2, p2=l1, power=(2.0, 2.0, 2.0)),
              light(_ffi.LIGHT_AREA, p0=l0, p1=l3, p2=l2, power=(2.0, 2.0, 2.0))]
    return CustomScene(triangles=tris, triangle_mat=[0] * 12 + [1, 2], materials=mats, lights=lights,
                       camera={"position": (0.5, 0.05, 0.5), "forward": (0.0, 1.0, 0.0), "res": (res, res)})


def _box_image(ctx, first, n, res=16, **params):
    sc = closed_white_box(res=res)
    sc.set_params(algorithm="pt", max_path_length=0, seed=17, **params)
    ctx.upload(sc)
    ctx.render(first, n)
    return ctx.framebuffer(0, sc.resolution)[0], sc


@pytest.mark.parametrize("params", [
    pytest.param(dict(splitting_budget=1.0, indirect_clipping=1e30), id="split_path"),
    pytest.param(dict(splitting_budget=1.2, indirect_clipping=1e30), id="tree_walker_chain")])
def test_deep_chain_fold_equals_the_forward_chain(ctx, params):
    """Chains of ~100 vertices, far beyond the 32 fold records a lane keeps: with a clipping
    threshold no estimate reaches, the split path (-sb 1, one sub-sample) and the tree walker's
    chain (-sb 1.2: one BSDF sample per vertex) compute the forward chain's estimator on the same
    RNG streams (branch 0), so their folds -- with every re-trace of the deep chains -- must give the
    chain kernel's image up to FP32 summation order."""
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)
    a, _ = _box_image(ctx, 0, 4, splitting_budget=1.0, indirect_clipping=0.0)
    b, _ = _box_image(ctx, 0, 4, **params)
    rel = np.abs(a - b) / np.maximum(np.abs(a), 1e-6)
    print(f"max per-channel relative difference {rel.max():.2e}")
    assert rel.max() < 1e-4
    assert abs(a.sum() - b.sum()) / a.sum() < 1e-6


@pytest.mark.parametrize("params", [
    pytest.param(dict(splitting_budget=4.0, indirect_clipping=16.0), id="split_first_vertex"),
    pytest.param(dict(splitting_budget=4.0, splitting_level=0.5, indirect_clipping=16.0), id="tree_walker")])
def test_deep_chains_agree_with_the_oracle(ctx, params):
    """The reference's defaults (splitting, clipping, Russian roulette) in the closed box: the GPU's
    mean radiance over 8192 samples per pixel against the oracle's over 512 others, each with
    its batch standard error (after ~100 bounces the two sides' paths part ways at ulp-level
    differences, so the same samples do not give the same pixels; their expectations must agree).
    A path cut at 48 vertices would lose 40 % of the box's radiance."""
    from oracle import oracle as O
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)
    g = []
    for b in range(32):
        img, sc = _box_image(ctx, b * 256, 256, res=8, **params)
        g.append(img.sum() / 256)
    orc = O.OracleScene(sc)
    o = [orc.render(1000000 + b * 32, 32, threads=0)[0].sum() / 32 for b in range(16)]
    g, o = np.array(g), np.array(o)
    se = np.sqrt(g.var(ddof=1) / len(g) + o.var(ddof=1) / len(o))
    z = (g.mean() - o.mean()) / se
    print(f"mean image sum gpu {g.mean():.6g} oracle {o.mean():.6g}: relative {g.mean() / o.mean() - 1:+.2e} "
          f"+- {se / o.mean():.1e} (z {z:+.2f})")
    assert abs(z) < 4.5
    assert se / o.mean() < 0.02


@pytest.mark.parametrize("scene_id", [3, 11, 18, 32, 34, 37])
def test_reference_defaults_on_reference_scenes(ctx, scene_id):
    """The reference's command-line defaults (-sb 4 -iic 16, Russian roulette: first-vertex
    splitting, SplitPath) on its scene IDs: Phong, GGX and smooth (Dirac) materials, area, point
    and background lights."""
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)
    sb = SceneBuild.reference(scene_id, resolution=24)
    sb.set_params(algorithm="pt", splitting_budget=4.0, indirect_clipping=16.0, max_path_length=0, seed=3)
    g, o = render_pair(ctx, sb, 4)
    assert_image_parity(g, o, frac=0.03)


@pytest.mark.parametrize("variant", [11, 40])
def test_several_lights_and_steerable_em_chain_parity(ctx, variant):
    """-sb 1 -iic 0 on a scene with three lights (two area triangles and a steerable EM): light
    picking by contribution, the EM as one of several lights in the shared EM lookups of escaped
    paths and EM light samples."""
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)
    sb = SceneBuild.workload(3, seed=2024, variant=variant, res=32)
    sb.set_params(seed=9)
    g, o = render_pair(ctx, sb, 4)
    assert_image_parity(g, o, frac=0.03)


@pytest.mark.parametrize("params", [dict(splitting_budget=4.0, indirect_clipping=16.0, max_path_length=0),
                                    dict(splitting_budget=3.0, indirect_clipping=2.0, max_path_length=6,
                                         light_to_bsdf_ratio=2.0)])
def test_splitting_with_several_lights_and_steerable_em(ctx, params):
    """Three lights (two area triangles and a steerable EM): light picking by contribution at the
    split vertex and along the chains, EM pdfs in the MIS of escaped subtrees."""
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)
    sb = SceneBuild.workload(3, seed=2024, variant=11, res=24)
    sb.set_params(algorithm="pt", seed=5, **params)
    g, o = render_pair(ctx, sb, 2)
    assert_image_parity(g, o, frac=0.03)


def test_splitting_on_reference_scene_with_em(ctx):
    """Reference scene 2 (box, spheres, EM background) with the reference's defaults -sb 4 -iic 16
    and Russian roulette."""
    ctx.set_kernel_mode(_ffi.KERNEL_WAVEFRONT)   # falls back to the tree kernel for these options
    sb = SceneBuild.reference(2, resolution=32)
    sb.set_params(algorithm="pt", splitting_budget=4.0, indirect_clipping=16.0, max_path_length=0)
    g, o = render_pair(ctx, sb, 4)
    assert_image_parity(g, o, frac=0.03)


def test_splitting_with_iic0_sb1_equals_chain_kernel(ctx):
    """-sb 1 -iic 0 takes the chain kernel; the tree walker reduces to it (same samples): render the
    chain kernel and the tree kernel (forced with a clipping threshold no radiance reaches)."""
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)
    sb = SceneBuild.workload(1, seed=33, res=32)
    ctx.upload(sb)
    ctx.render(0, 4)
    a = ctx.framebuffer(0, sb.resolution)[0]
    sb.set_params(indirect_clipping=1e30)
    ctx.upload(sb)
    ctx.render(0, 4)
    b = ctx.framebuffer(0, sb.resolution)[0]
    rel = np.abs(a - b) / np.maximum(np.abs(a), 1e-6)
    assert np.mean(rel > 1e-5) < 0.02
    assert abs(a.sum() - b.sum()) / a.sum() < 1e-4


@pytest.mark.parametrize("kernel,params", [
    (_ffi.KERNEL_MEGAKERNEL, {}), (_ffi.KERNEL_WAVEFRONT, {}),
    (_ffi.KERNEL_MEGAKERNEL, dict(splitting_budget=4.0, indirect_clipping=16.0, max_path_length=0)),
    (_ffi.KERNEL_MEGAKERNEL, dict(algorithm="ptn"))], ids=["chain", "wavefront", "split_rr", "ptn"])
def test_single_pixel_image(ctx, kernel, params):
    """A 1x1 image (one partly filled 8x8 tile): every kernel organisation against the oracle,
    over enough iterations that the pixel's sum is a stable comparison."""
    ctx.set_kernel_mode(kernel)
    sb = SceneBuild.workload(1, seed=31, res=1)
    sb.set_params(**params)
    g, o = render_pair(ctx, sb, 256)
    assert g.shape == (1, 1, 3) and np.isfinite(g).all() and g.sum() > 0
    assert abs(g.sum() - o.sum()) / o.sum() < 1e-3
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)


def test_empty_iteration_range_is_a_no_op(ctx):
    for k in (_ffi.KERNEL_MEGAKERNEL, _ffi.KERNEL_WAVEFRONT):
        ctx.set_kernel_mode(k)
        sb = SceneBuild.workload(1, seed=32, res=16)
        ctx.upload(sb)
        ctx.render(0, 0)
        g, it = ctx.framebuffer(0, sb.resolution)
        assert it == 0 and not g.any()
        ctx.render(0, 2)
        a, _ = ctx.framebuffer(0, sb.resolution)
        ctx.render(2, 0)
        b, it = ctx.framebuffer(0, sb.resolution)
        assert it == 2 and np.array_equal(a, b)
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)


@pytest.mark.parametrize("min_pl,max_pl", [(1, 1), (2, 2), (2, 8), (3, 5)])
@pytest.mark.parametrize("kernel", KERNELS)
def test_path_length_bounds(ctx, kernel, min_pl, max_pl):
    """-minpl / -maxpl (path_tracer.hxx:198-210, 252-265): direct light only, one bounce only,
    emission from the camera ray skipped."""
    ctx.set_kernel_mode(kernel)
    sb = SceneBuild.workload(2, seed=33, res=24)
    sb.set_params(min_path_length=min_pl, max_path_length=max_pl)
    g, o = render_pair(ctx, sb, 4)
    assert_image_parity(g, o, frac=0.03)
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)


# ---- scene intersection (scene.hxx:42-74), function level ---------------------------------------
from geom64 import cast as cast64, random_rays  # noqa: E402

ISECT_SCENES = [pytest.param(lambda: SceneBuild.workload(1, seed=51, res=8), id="workload1"),
                pytest.param(lambda: SceneBuild.workload(2, seed=52, res=8), id="workload2"),
                pytest.param(lambda: SceneBuild.reference(2, em_type=0, resolution=8), id="reference_scene2"),
                pytest.param(lambda: SceneBuild.reference(11, em_type=0, resolution=8), id="reference_scene11")]


def _isect_rays(seed, n=100000):
    rng = np.random.default_rng(seed)
    org_in, dir_in = random_rays(rng, n // 2)
    org_out, dir_out = random_rays(rng, n - n // 2, -1.0, 2.0)
    org, dir = np.concatenate([org_in, org_out]), np.concatenate([dir_in, dir_out])
    tmin = rng.choice([0.0, 1e-3, 0.05], len(org)).astype(np.float32)
    return rng, org, dir, tmin


@pytest.mark.parametrize("build", ISECT_SCENES)
def test_intersect_parity(ctx, build):
    """pg3g_intersect_batch against the oracle on 100k rays from inside and outside the scene:
    the same hit/miss and material, distances and normals to float32 rounding (the GPU tests
    triangle edges as Pluecker lines and contracts FMAs), except on rays within 1e-5 of an edge,
    a tangent, a bound or a second surface (classified by an independent FP64 ray caster)."""
    sb = build()
    ctx.upload(sb)
    _, org, dir, tmin = _isect_rays(61)
    dg, ng, mg = ctx.intersect(org, dir, tmin)
    do, no, mo = O.OracleScene(sb).intersect(org, dir, tmin)
    _, _, _, margin = cast64(sb, org, dir, tmin)
    clear = margin > 1e-5
    assert clear.mean() > 0.99
    assert np.array_equal((dg >= 0)[clear], (do >= 0)[clear])
    assert np.array_equal(mg[clear], mo[clear])
    both = clear & (do >= 0)
    assert both.mean() > 0.3
    rel = np.abs(dg[both] - do[both]) / np.maximum(do[both], 1e-2)
    assert rel.max() < 1e-5, rel.max()
    # normals: near-tangent sphere hits amplify an ulp of the quadratic's coefficients by
    # 1/sqrt(discriminant); the bulk agrees to float32 rounding
    en = np.abs(ng[both] - no[both]).max(axis=1)
    assert en.max() < 1e-4 and np.quantile(en, 0.999) < 2e-6, (en.max(), np.quantile(en, 0.999))
    # all rays, ambiguous ones included: disagreements stay rare
    assert np.mean((dg >= 0) != (do >= 0)) < 1e-3


@pytest.mark.parametrize("build", ISECT_SCENES)
def test_occluded_parity(ctx, build):
    sb = build()
    ctx.upload(sb)
    rng, p, dir, _ = _isect_rays(62)
    tmax = rng.uniform(0.05, 1.5, len(p)).astype(np.float32)
    og = ctx.occluded(p, dir, tmax)
    oo = O.OracleScene(sb).occluded(p, dir, tmax)
    org = p.astype(np.float64) + 1e-3 * dir.astype(np.float64)
    _, _, _, margin = cast64(sb, org, dir, 0.0, tmax.astype(np.float64) - 2e-3)
    clear = margin > 1e-5
    assert clear.mean() > 0.99
    assert np.array_equal(og[clear], oo[clear])
    assert np.mean(og != oo) < 1e-3
    assert 0.05 < oo.mean() < 0.95


@pytest.mark.parametrize("build", ISECT_SCENES)
def test_rectangles_equal_their_triangles(ctx, build, monkeypatch, request):
    """The walls merged into axis-aligned rectangles at upload (pairs of triangles tested as one
    primitive) hit exactly as their two triangles: on 100k rays the same hit/miss and material and
    bit-identical distances and normals, except on rays within 1e-5 of an edge, a tangent or a
    second surface (independent FP64 caster); Scene::Occluded likewise."""
    sb = build()
    rng, org, dir, tmin = _isect_rays(63)
    tmax = rng.uniform(0.05, 1.5, len(org)).astype(np.float32)
    ctx.upload(sb)
    rect, tris, sphs = ctx.primitives()
    dq, nq, mq = ctx.intersect(org, dir, tmin)
    oq = ctx.occluded(org, dir, tmax)
    monkeypatch.setenv("PG3G_TRIANGLES_ONLY", "1")
    ctx.upload(sb)
    monkeypatch.delenv("PG3G_TRIANGLES_ONLY")
    assert ctx.primitives() == (0, tris + 2 * rect, sphs)
    dt, nt, mt = ctx.intersect(org, dir, tmin)
    ot = ctx.occluded(org, dir, tmax)
    ctx.upload(sb)
    print(f"  {rect} rectangles, {tris} triangles, {sphs} spheres")
    if "workload" in request.node.name:
        assert rect == 5   # the unit box's five walls; the reference boxes' halves differ in the last ulp of n
    _, _, _, margin = cast64(sb, org, dir, tmin)
    clear = margin > 1e-5
    assert clear.mean() > 0.99
    assert np.array_equal(mq[clear], mt[clear])
    assert np.array_equal(dq[clear].view(np.uint32), dt[clear].view(np.uint32))
    assert np.array_equal(nq[clear].view(np.uint32), nt[clear].view(np.uint32))
    assert np.mean(mq != mt) < 1e-3
    p64 = org.astype(np.float64) + 1e-3 * dir.astype(np.float64)
    _, _, _, margin_o = cast64(sb, p64, dir, 0.0, tmax.astype(np.float64) - 2e-3)
    clear_o = margin_o > 1e-5
    assert np.array_equal(oq[clear_o], ot[clear_o])
    assert np.mean(oq != ot) < 1e-3


ALL_SCENE_IDS = [i for i in range(38) if i != 27]   # 27: the SRA layered material, outside the GPU path


@pytest.mark.parametrize("mode", ["chain", "chain_wavefront", "defaults"])
@pytest.mark.parametrize("scene_id", ALL_SCENE_IDS)
def test_every_reference_scene_id(ctx, scene_id, mode):
    """Every reference scene ID the GPU path supports (config.hxx:45-115, scene.hxx:176-956),
    with the chain settings (-sb 1 -iic 0 -maxpl 8; fused kernel and wavefront pipeline) and with
    the reference's command-line defaults (-sb 4 -iic 16, Russian roulette)."""
    ctx.set_kernel_mode(_ffi.KERNEL_WAVEFRONT if mode == "chain_wavefront" else _ffi.KERNEL_MEGAKERNEL)
    sb = SceneBuild.reference(scene_id, resolution=24)
    if mode.startswith("chain"):
        sb.set_params(algorithm="pt", splitting_budget=1.0, indirect_clipping=0.0, max_path_length=8, seed=9)
    else:
        sb.set_params(algorithm="pt", splitting_budget=4.0, indirect_clipping=16.0, max_path_length=0, seed=9)
    g, o = render_pair(ctx, sb, 4)
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)
    assert_image_parity(g, o, frac=0.03)


@pytest.mark.parametrize("alg", ["ptn", "dbs", "dlsa", "dlss", "dmis"])
@pytest.mark.parametrize("scene_id", ALL_SCENE_IDS)
def test_every_reference_scene_id_other_algorithms(ctx, scene_id, alg):
    """The naive path tracer and the four direct-illumination estimators (path_tracer.hxx:45-165,
    direct_illumination.hxx:15-156) on every supported reference scene ID."""
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)
    sb = SceneBuild.reference(scene_id, resolution=24)
    sb.set_params(algorithm=alg, max_path_length=8, seed=10)
    g, o = render_pair(ctx, sb, 4)
    if o.sum() == 0:      # e.g. point lights only, which ptn and dbs cannot hit
        assert g.sum() == 0
        return
    assert_image_parity(g, o, frac=0.03)


EM_SCENE_IDS = [6, 7, 8, 9, 10, 12, 13, 14, 15, 20, 21, 22, 23, 24, 25, 26, 28, 29, 37]   # kLightEnv scenes


@pytest.mark.parametrize("sampler", [_ffi.EM_SIMPLE, _ffi.EM_STEERABLE], ids=["simple", "steerable"])
@pytest.mark.parametrize("mode", ["chain", "defaults"])
@pytest.mark.parametrize("scene_id", EM_SCENE_IDS)
def test_every_em_scene_with_image_em(ctx, scene_id, mode, sampler, tmp_path):
    """The reference's environment-light scene IDs with an image EM read from a file (em type 7,
    rotation 0.83 and scale 1.5, scene.hxx:891-897): a 128x64 probe of a sky gradient and three
    bright lobes, under both EM samplers, with the chain settings and with the reference's
    defaults (-sb 4 -iic 16, Russian roulette)."""
    ctx.set_kernel_mode(_ffi.KERNEL_MEGAKERNEL)
    yy, xx = np.meshgrid(np.arange(64), np.arange(128), indexing="ij")
    img = 0.2 + 0.8 * (1 - yy / 64)[..., None] * np.array([0.5, 0.75, 1.0])
    for cx, cy, peak in ((30, 12, 60.0), (90, 20, 25.0), (64, 40, 10.0)):
        img = img + peak * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / 18.0)[..., None]
    path = str(tmp_path / "probe.pfm")
    with open(path, "wb") as f:
        f.write(b"PF\n128 64\n-1.0\n" + np.ascontiguousarray(img[::-1]).astype("<f4").tobytes())
    sb = SceneBuild.reference(scene_id, em_type=7, resolution=24, em_path=path)
    sb.desc.em.sampler = sampler
    if mode == "chain":
        sb.set_params(algorithm="pt", splitting_budget=1.0, indirect_clipping=0.0, max_path_length=8, seed=11)
    else:
        sb.set_params(algorithm="pt", splitting_budget=4.0, indirect_clipping=16.0, max_path_length=0, seed=11)
    g, o = render_pair(ctx, sb, 4)
    assert_image_parity(g, o, frac=0.03)
