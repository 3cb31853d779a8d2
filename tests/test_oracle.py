"""Pins the CPU oracle (oracle/) before it is trusted as the checker of the GPU path.

Each test checks an oracle function against the reference's own known-answer tests or against
what the mathematics fixes (closed forms, invariants, brute force), chosen so that a dropped term,
wrong sign or transposed operand fails.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from pg3render_amd import SceneBuild, _ffi
from pg3render_amd.scene_desc import CustomScene, light, material

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def direction(theta, phi):   # Geom::CreateDirection, geom.hxx:30-61
    return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)], np.float32)


# ---------------------------------------------------------------------------------------------
# RNG: Philox4x32-10 known-answer vectors of Random123 (Salmon et al., SC'11; kat_vectors)
@pytest.mark.parametrize("ctr,key,expected", [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox_known_answers(ctr, key, expected):
    seed = key[0] | (key[1] << 32)
    assert tuple(int(v) for v in O.philox(*ctr, seed)) == expected


# ---------------------------------------------------------------------------------------------
# Steerable SH basis: the reference's 18 canonical/diagonal KATs (em_steerable_sampler.hxx:276-646)
def test_sh_basis_reference_kats():
    kat = json.load(open(os.path.join(GOLDEN, "sh_basis_kats.json")))
    c = math.cos(math.pi / 4)
    sym = {"c": c, "y20": 0.5 * (3 * c * c - 1), "c2": c * c}
    norm = np.array(kat["normalization"])

    def val(s):
        neg = s.startswith("-")
        s = s.lstrip("-")
        v = sym[s] if s in sym else float(s)
        return -v if neg else v

    for case in kat["cases"]:
        d = direction(case["theta_pi"] * math.pi, case["phi_pi"] * math.pi)
        expected = np.array([val(s) for s in case["normalized"]]) * norm
        got = O.sh_basis(d)
        assert np.max(np.abs(got - expected)) <= 1e-4, case["name"]


@pytest.mark.parametrize("compensate", [False, True])
def test_clamped_cosine_reference_bounds(compensate):
    """_UT_GenerateForClampedCos bounds (em_steerable_sampler.hxx:734-849) on 200 x 2000 directions."""
    b = json.load(open(os.path.join(GOLDEN, "sh_basis_kats.json")))["clamped_cos_bounds"]
    b = b["compensated" if compensate else "uncompensated"]
    rng = np.random.default_rng(5)

    def sphere(n):
        u = rng.random((n, 2))
        z = 1 - 2 * u[:, 1]
        r = np.sqrt(np.maximum(0, 1 - z * z))
        return np.stack([r * np.cos(2 * np.pi * u[:, 0]), r * np.sin(2 * np.pi * u[:, 0]), z], 1).astype(np.float32)

    normals, dirs = sphere(200), sphere(2000)
    basis = np.stack([O.sh_basis(d) for d in dirs])
    for n in normals:
        coeff = O.clamped_cos(n, compensate)
        recon = basis @ coeff
        direct = np.maximum(dirs @ n, 0)
        assert recon.max() <= b["max"]
        assert recon[direct == 0].max() <= b["max_in_zero_area"]
        assert recon.min() >= b["min"] - 1e-6
        diff = recon - direct
        assert diff.min() >= b["diff_min"] - 1e-6 and diff.max() <= b["diff_max"]


def test_clamped_cosine_projection_closed_form():
    """Without compensation the coefficients are the SH projection of max(cos,0) (Ramamoorthi-Hanrahan):
    the reconstruction at the normal is pi*(1/4 + 1/2 + 5/16)/pi ... i.e. c4*Y00 + 2c2*Y10 + ... = 1.0217."""
    n = np.array([0, 0, 1], np.float32)
    coeff = O.clamped_cos(n, False)
    at_normal = float(O.sh_basis(n) @ coeff)
    # truncated expansion of the clamped cosine at theta = 0: 1/4 + 1/2 + 5/16 (orders 0, 1, 2)
    assert abs(at_normal - (0.25 + 0.5 + 0.3125)) < 1e-3
    # opposite direction: 1/4 - 1/2 + 5/16
    assert abs(float(O.sh_basis(-n) @ coeff) - (0.25 - 0.5 + 0.3125)) < 1e-3


# ---------------------------------------------------------------------------------------------
# Fresnel (physics.hxx)
@pytest.mark.parametrize("eta,k", [(0.2, 3.8), (0.468, 2.81), (0.121, 3.66), (1.5, 0.5)])
def test_fresnel_conductor_normal_incidence(eta, k):
    f0 = ((eta - 1) ** 2 + k * k) / ((eta + 1) ** 2 + k * k)
    assert abs(O.fresnel_conductor(1.0, eta, k) - f0) < 1e-6
    assert O.fresnel_conductor(-0.1, eta, k) == 0.0
    # grazing incidence reflects everything
    assert abs(O.fresnel_conductor(0.0, eta, k) - 1.0) < 1e-5


@pytest.mark.parametrize("eta", [1.5, 1.33, 2.4])
def test_fresnel_dielectric_closed_forms(eta):
    assert abs(O.fresnel_dielectric(1.0, eta) - ((eta - 1) / (eta + 1)) ** 2) < 1e-6
    # from inside beyond the critical angle: total internal reflection
    crit = math.sqrt(1 - 1 / eta ** 2)
    assert O.fresnel_dielectric(-(crit - 0.05), eta) == 1.0
    # Stokes symmetry: F(theta_i outside) == F(theta_t inside)
    for ci in (0.9, 0.6, 0.3, 0.1):
        st = math.sqrt(1 - ci * ci) / eta
        ct = math.sqrt(1 - st * st)
        assert abs(O.fresnel_dielectric(ci, eta) - O.fresnel_dielectric(-ct, eta)) < 1e-5


# ---------------------------------------------------------------------------------------------
# GGX (microfacet.hxx)
@pytest.mark.parametrize("alpha", [0.05, 0.1, 0.3, 1.0])
def test_ggx_distribution_normalised(alpha):
    """int D(h) cos(theta_h) dw = 1."""
    th = (np.arange(20000) + 0.5) / 20000 * (math.pi / 2)
    vals = np.array([O.ggx_d(direction(t, 0.3), alpha) for t in th])
    integral = np.sum(vals * np.cos(th) * np.sin(th)) * (math.pi / 2) / 20000 * 2 * math.pi
    assert abs(integral - 1.0) < 2e-3


@pytest.mark.parametrize("alpha", [0.1, 0.5])
def test_smith_g1_closed_form(alpha):
    for theta in (0.1, 0.7, 1.3):
        v = direction(theta, 0.4)
        g = O.ggx_g1(v, np.array([0, 0, 1], np.float32), alpha)
        assert abs(g - 2 / (1 + math.sqrt(1 + alpha ** 2 * math.tan(theta) ** 2))) < 1e-6
    # back-facing microfacet or direction below the facet: zero
    assert O.ggx_g1(direction(0.3, 0), direction(2.0, 0), alpha) == 0.0


@pytest.mark.parametrize("alpha,theta_o", [(0.1, 0.2), (0.3, 1.0), (0.6, 1.4)])
def test_vndf_pdf_integrates_to_one(alpha, theta_o):
    """D_wo(h) = G1(wo,h) max(0, wo.h) D(h) / wo.z integrates to 1 over h (Heitz 2014)."""
    wo = direction(theta_o, 0.0)
    nt, npz = 400, 200
    th = (np.arange(nt) + 0.5) / nt * (math.pi / 2)
    ph = (np.arange(npz) + 0.5) / npz * 2 * math.pi
    tot = 0.0
    for t in th:
        s = 0.0
        for p in ph[::4]:
            s += O.vndf_pdf(wo, direction(t, p), alpha)
        tot += s * 4 * math.sin(t)
    tot *= (math.pi / 2) / nt * 2 * math.pi / npz
    assert abs(tot - 1.0) < 0.02


def test_vndf_sampling_matches_pdf_moments():
    """Sampled visible normals have the moments of D_wo (chi-square style check on h.z and h.x)."""
    alpha, wo = 0.4, direction(0.9, 0.0)
    rng = np.random.default_rng(3)
    u = rng.random((20000, 2)).astype(np.float32)
    hs = np.array([O.vndf_sample(wo, alpha, a, b) for a, b in u])
    assert np.all(hs[:, 2] > 0)
    assert np.allclose(np.linalg.norm(hs, axis=1), 1, atol=1e-5)
    nt, npz = 200, 64
    th = (np.arange(nt) + 0.5) / nt * (math.pi / 2)
    ph = (np.arange(npz) + 0.5) / npz * 2 * math.pi
    m_z = m_x = norm = 0.0
    for t in th:
        for p in ph:
            h = direction(t, p)
            w = O.vndf_pdf(wo, h, alpha) * math.sin(t)
            norm += w
            m_z += w * h[2]
            m_x += w * h[0]
    assert abs(hs[:, 2].mean() - m_z / norm) < 0.01
    assert abs(hs[:, 0].mean() - m_x / norm) < 0.01


def test_halfway_refraction_reference_consistency():
    """_UT_HalfwayVectorRefractionLocal (microfacet.hxx:125-501): for in/out directions on opposite
    sides, the refraction halfway vector refracts the input into the output."""
    checked = 0
    for upper, lower in ((1.0, 1.51), (1.1, 1.0), (1.51, 1.55)):
        for ti in np.linspace(0.05, math.pi - 0.05, 16):
            for to in np.linspace(0.05, math.pi - 0.05, 16):
                for po in np.linspace(0, 2 * math.pi, 8, endpoint=False):
                    di, do = direction(ti, 0.0), direction(to, po)
                    if (di[2] < 0) == (do[2] < 0):
                        continue
                    eta_io = (lower if do[2] < 0 else upper) / (lower if di[2] < 0 else upper)
                    h = O.halfway_refraction(di, do, eta_io)
                    cim, com = float(di @ h), float(do @ h)
                    if not (cim * com <= 0 and di[2] * cim >= 0 and do[2] * com >= 0):
                        continue
                    eta = lower / upper
                    ei = 1 / eta if cim > 0 else eta
                    if 1 - (1 - cim * cim) * ei * ei < 0:
                        continue
                    out = O.refract(di, h, eta)
                    assert np.linalg.norm(out - do) < 5e-4
                    checked += 1
    assert checked > 500


# ---------------------------------------------------------------------------------------------
# BSDFs: closed forms and sample/eval/pdf consistency (materials.hxx)
def _mat(kind, **kw):
    return material(kind, **kw)


def _hemi(rng, n, lower=False):
    u = rng.random((n, 2))
    z = np.sqrt(u[:, 1]) * 0.98 + 0.01
    r = np.sqrt(1 - z * z)
    d = np.stack([r * np.cos(2 * np.pi * u[:, 0]), r * np.sin(2 * np.pi * u[:, 0]), z], 1)
    if lower:
        d[:, 2] *= -1
    return d.astype(np.float32)


def test_lambert_closed_form():
    rng = np.random.default_rng(1)
    m = _mat(_ffi.MAT_LAMBERT, diffuse=(0.8, 0.1, 0.3))
    wi, wo = _hemi(rng, 64), _hemi(rng, 64)
    f, pdf, comp = O.bsdf_eval(m, wi, wo)
    assert np.allclose(f, np.array([0.8, 0.1, 0.3]) / math.pi, rtol=1e-6)
    assert np.allclose(pdf, wi[:, 2] / math.pi, rtol=1e-6)
    f2, _, _ = O.bsdf_eval(m, _hemi(rng, 8, lower=True), _hemi(rng, 8))
    assert np.all(f2 == 0)


@pytest.mark.parametrize("kind,kw", [
    (_ffi.MAT_LAMBERT, dict(diffuse=(0.6, 0.6, 0.6))),
    (_ffi.MAT_GGX_CONDUCTOR, dict(alpha=0.1, inner_ior=0.2, absorbance=3.8)),
    (_ffi.MAT_GGX_CONDUCTOR, dict(alpha=0.5, inner_ior=0.468, absorbance=2.81)),
])
def test_reciprocity(kind, kw):
    rng = np.random.default_rng(2)
    m = _mat(kind, **kw)
    wi, wo = _hemi(rng, 256), _hemi(rng, 256)
    f1, _, _ = O.bsdf_eval(m, wi, wo)
    f2, _, _ = O.bsdf_eval(m, wo, wi)
    assert np.allclose(f1, f2, rtol=2e-5, atol=1e-7)


@pytest.mark.parametrize("kind,kw", [
    (_ffi.MAT_GGX_CONDUCTOR, dict(alpha=0.1, inner_ior=0.2, absorbance=3.8)),
    (_ffi.MAT_GGX_CONDUCTOR, dict(alpha=0.45, inner_ior=0.468, absorbance=2.81)),
    (_ffi.MAT_GGX_DIELECTRIC, dict(alpha=0.3, inner_ior=1.5)),
    (_ffi.MAT_GGX_DIELECTRIC, dict(alpha=0.05, inner_ior=1.33)),
])
def test_ggx_sample_weight_closed_form(kind, kw):
    """VNDF sampling weight f*|cos_i|/pdf equals F(o.h) G1(i,h) for the conductor and G1(i,h) for
    the dielectric (reflection and refraction, Heitz & d'Eon 2014); eval at the sampled direction
    reproduces the sampled value and pdf."""
    rng = np.random.default_rng(4)
    m = _mat(kind, **kw)
    n = 2000
    wo = _hemi(rng, n, lower=False)
    if kind == _ffi.MAT_GGX_DIELECTRIC:
        wo[n // 2:, 2] *= -1    # both sides
    u = rng.random((n, 3)).astype(np.float32)
    wil, f, pdf, comp = O.bsdf_sample(m, wo, u)
    fe, pe, ce = O.bsdf_eval(m, wil, wo)
    good = (f[:, 0] > 0) & (pdf > 1e-3) & (np.abs(wil[:, 2]) > 0.02) & (np.abs(wo[:, 2]) > 0.02)
    assert good.sum() > n * 0.5
    alpha = kw["alpha"]
    checked = 0
    for i in np.nonzero(good)[0][:400]:
        o, w = wo[i], wil[i]
        s = -1.0 if o[2] < 0 else 1.0
        oS, iS = o * s, w * s
        refl = (iS[2] > 0)
        if refl:
            h = (oS + iS) / np.linalg.norm(oS + iS)
        else:
            eta = kw["inner_ior"] if s > 0 else 1 / kw["inner_ior"]
            h = -(iS * eta + oS)
            h = h / np.linalg.norm(h)
            h = h if h[2] > 0 else -h
        g1 = O.ggx_g1(iS, h, alpha)
        weight = f[i, 0] * abs(w[2]) / pdf[i]
        expected = g1 * (O.fresnel_conductor(float(o @ h), kw["inner_ior"], kw["absorbance"])
                         if kind == _ffi.MAT_GGX_CONDUCTOR else 1.0)
        assert abs(weight - expected) <= 2e-3 * max(1.0, expected), (i, weight, expected)
        assert abs(fe[i, 0] - f[i, 0]) <= 2e-3 * f[i, 0]
        assert abs(pe[i] - pdf[i]) <= 2e-3 * pdf[i]
        checked += 1
    assert checked >= 200


def test_cosine_and_uniform_triangle_warps():
    """Lambert sampling is cos-distributed (E[z] = 2/3); pdf = z/pi at the sample."""
    rng = np.random.default_rng(6)
    m = _mat(_ffi.MAT_LAMBERT, diffuse=(0.5, 0.5, 0.5))
    u = rng.random((50000, 3)).astype(np.float32)
    wil, f, pdf, comp = O.bsdf_sample(m, np.tile([[0, 0, 1]], (50000, 1)), u)
    assert abs(wil[:, 2].mean() - 2 / 3) < 5e-3
    assert np.allclose(pdf, wil[:, 2] / math.pi, rtol=1e-5)


# ---------------------------------------------------------------------------------------------
# Distributions (distribution.hxx)
def test_distribution1d_brute_force():
    rng = np.random.default_rng(7)
    f = rng.random(37).astype(np.float32)
    f[5] = 0.0
    u = rng.random(5000).astype(np.float32)
    x, seg, pdf, cdf, cdf32 = O.dist1d(f, u)
    ref = np.concatenate([[0.0], np.cumsum(f.astype(np.float64) / len(f))])
    ref /= ref[-1]
    assert np.allclose(cdf, ref, atol=2e-7)
    assert np.allclose(cdf32, ref, atol=2e-6)          # the reference's FP32 accumulation
    ut = u * np.float32(0.999999)
    assert np.all((ut >= cdf[seg]) & (ut < cdf[seg + 1]))
    assert np.all(seg != 5)                             # zero segment never sampled
    assert np.allclose(pdf, (cdf[seg + 1] - cdf[seg]) * len(f), rtol=1e-6)
    assert np.all((x >= seg / len(f)) & (x <= (seg + 1) / len(f)))


def test_distribution1d_zero_function_is_uniform():
    f = np.zeros(8, np.float32)
    x, seg, pdf, cdf, _ = O.dist1d(f, np.array([0.0, 0.3, 0.99], np.float32))
    assert np.allclose(cdf, np.arange(9) / 8)
    assert np.allclose(pdf, 1.0)


def test_distribution_fp32_vs_fp64_accumulation_bound():
    rng = np.random.default_rng(8)
    f = (rng.random(1024) ** 8 * 1e4).astype(np.float32)
    _, _, _, cdf, cdf32 = O.dist1d(f, np.zeros(1, np.float32))
    assert np.max(np.abs(cdf - cdf32)) < 1e-5


# ---------------------------------------------------------------------------------------------
# EM image and samplers (em_image.hxx, filter.hxx, em_simple_spherical_sampler.hxx)
def _em_scene(em, sampler=_ffi.EM_SIMPLE, levels=(4, 4)):
    return CustomScene(materials=[material(_ffi.MAT_LAMBERT, diffuse=(0.5, 0.5, 0.5))],
                       lights=[light(_ffi.LIGHT_BACKGROUND)], em=em, em_sampler=sampler,
                       steerable_levels=levels, camera={"res": (8, 8)})


def test_tent_integral_reference_kats():
    """Filter::TriangleIntegral KATs (filter.hxx:123-173) through AveragePixelLuminance."""
    kats = json.load(open(os.path.join(GOLDEN, "sh_basis_kats.json")))["tent_integral_kats"]["cases"]
    for case in kats:
        g = np.array(case["grid"], np.float32)
        img = np.zeros((6, 8, 3), np.float32)
        img[1:4, 2:5, :] = g.T[:, :, None]   # grid[x][y] -> image row y, column x
        sc = _em_scene(img)
        o = O.OracleScene(sc)
        got = O.load().oracle_em_avg_pixel_lum(o.ptr, 3, 2)
        assert abs(got - case["expected"]) < 1e-4


def test_em_tent_filter_interpolates_texel_centres():
    rng = np.random.default_rng(9)
    img = rng.random((16, 32, 3)).astype(np.float32)
    o = O.OracleScene(_em_scene(img))
    xs, ys = np.meshgrid(np.arange(32), np.arange(1, 15))
    uv = np.stack([(xs.ravel() + 0.5) / 32, (ys.ravel() + 0.5) / 16], 1).astype(np.float32)
    got = o.em_eval_uv(uv)
    assert np.allclose(got, img[ys.ravel(), xs.ravel()], atol=1e-5)


def test_simple_sampler_pdf_normalised_and_consistent():
    rng = np.random.default_rng(10)
    img = (0.2 + rng.random((32, 64, 3)) ** 4 * 10).astype(np.float32)
    o = O.OracleScene(_em_scene(img))
    n = 200000
    u = rng.random((n, 2))
    z = 1 - 2 * u[:, 1]
    r = np.sqrt(1 - z * z)
    d = np.stack([r * np.cos(2 * np.pi * u[:, 0]), r * np.sin(2 * np.pi * u[:, 0]), z], 1).astype(np.float32)
    normals = np.tile([[0, 0, 1]], (n, 1)).astype(np.float32)
    pdf = o.em_pdf(d, normals, np.ones(n, np.int32))
    assert abs(pdf.mean() * 4 * math.pi - 1.0) < 0.02
    us = rng.random((5000, 3)).astype(np.float32)
    sd, sp, rc, ok = o.em_sample(normals[:5000], np.ones(5000, np.int32), us)
    assert np.all(ok == 1)
    pe = o.em_pdf(sd, normals[:5000], np.ones(5000, np.int32))
    close = np.abs(pe - sp) <= 1e-4 * sp
    assert close.mean() > 0.97      # equal except where Dir2LatLong lands in a neighbouring segment


def test_steerable_tree_reference_invariants():
    """_UT_Init / _UT_BuildTriangleTreeSynthetic (em_steerable_sampler.hxx:4518-4775): a constant map
    refines to 20*4^L leaves, the tree has ceil(log6 n)+1 levels and the root weight is the sum of
    the leaf weights; leaves are unit-sphere triangles covering the sphere."""
    img = np.ones((16, 32, 3), np.float32)
    for lvl in (0, 1, 2, 3):
        o = O.OracleScene(_em_scene(img, _ffi.EM_STEERABLE, (lvl, lvl)))
        leaves, nverts, w = o.steerable()
        assert len(leaves) == 20 * 4 ** lvl
        assert np.allclose(np.linalg.norm(leaves, axis=2), 1, atol=1e-5)
        area = 0.5 * np.linalg.norm(np.cross(leaves[:, 1] - leaves[:, 0], leaves[:, 2] - leaves[:, 1]), axis=1)
        # constant unit luminance: w[0] = 0.282095 * sum(planar areas), w[1..3] ~ 0 by symmetry
        assert abs(w[0] - 0.282095 * area.sum()) < 1e-3 * w[0]
        assert np.all(np.abs(w[1:4]) < 1e-3 * w[0])


def test_steerable_pick_frequencies():
    """_UT_Sampling_EM (em_steerable_sampler.hxx:4778-4925): triangle picking frequency matches the
    relative integral within 0.01."""
    rng = np.random.default_rng(11)
    img = (0.1 + rng.random((16, 32, 3)) ** 3).astype(np.float32)
    o = O.OracleScene(_em_scene(img, _ffi.EM_STEERABLE, (1, 1)))
    for n in ([0, 0, 1], [0.6, 0.0, 0.8], [-0.48, 0.6, -0.64]):
        u = rng.random(200000).astype(np.float32)
        leaf, rel = o.steerable_pick(np.array(n, np.float32), u)
        freq = np.bincount(leaf, minlength=len(rel)) / len(u)
        assert np.max(np.abs(freq - rel)) < 0.01


def test_triangle_bilinear_marginal_inverts_linear_density():
    """SampleTriangleFFunction (em_steerable_sampler.hxx:3709-3724) inverts the marginal CDF of the
    linear density v(p) = b0*a + b1*b + b2*c over the triangle: with s = 1 - b0 and the area element
    s ds dt, F(s) = (y s^2 + x s^3) / (x + y), x = (b + c - 2a)/6, y = a/2.  Checked by a KS
    statistic.  The conditional part (SampleTriangleGFunction, :3728-3747) is parity unpinned: its
    normaliser omits an s*b term, so it is not the inverse conditional CDF (the reference's own
    single-triangle test only reports this, :5153-5161, and Debugging/*.lyx documents it)."""
    rng = np.random.default_rng(12)
    for a, b, c in ((1.0, 1.0, 1.0), (0.2, 1.0, 3.0), (2.0, 0.5, 0.1)):
        u = rng.random((20000, 2))
        bary = np.array([O.sample_tri_bilinear(float(p), float(q), a, b, c) for p, q in u])
        assert np.all(bary >= -1e-6) and np.all(bary.sum(1) <= 1 + 1e-5)
        s = np.sort(1.0 - bary[:, 0])
        x, y = (b + c - 2 * a) / 6, a / 2
        cdf = (y * s ** 2 + x * s ** 3) / (x + y)
        ks = np.max(np.abs(cdf - (np.arange(1, len(s) + 1) / len(s))))
        assert ks < 0.015, (a, b, c, ks)


# ---------------------------------------------------------------------------------------------
# Camera and geometry
def test_camera_centre_ray_and_fov():
    sb = SceneBuild.workload(1, seed=1, res=64)
    o = O.OracleScene(sb)
    org, d = o.camera_ray(32.0, 32.0)
    assert np.allclose(org, [0.5, -1.25, 0.5])
    assert np.allclose(d, [0, 1, 0], atol=1e-6)
    _, dl = o.camera_ray(0.0, 32.0)
    _, dr = o.camera_ray(64.0, 32.0)
    ang = math.degrees(math.acos(float(np.clip(dl @ dr, -1, 1))))
    assert abs(ang - 45.0) < 1e-3


# ---------------------------------------------------------------------------------------------
# Image-level invariants of the oracle's estimators
def _box_with(algorithm, res=32, maxpl=4, light_kind="area"):
    sb = SceneBuild.workload(1, seed=3, res=res)
    sb.set_params(algorithm=algorithm, max_path_length=maxpl)
    return sb


def test_white_furnace_lambert_sphere():
    """A convex Lambert sphere (albedo 0.5) under constant unit radiance reflects exactly 0.5."""
    sc = CustomScene(spheres=[[0, 0, 0, 1]], sphere_mat=[0],
                     materials=[material(_ffi.MAT_LAMBERT, diffuse=(0.5, 0.5, 0.5))],
                     lights=[light(_ffi.LIGHT_BACKGROUND, power=(1, 1, 1))],
                     camera={"position": (0, -4, 0), "forward": (0, 1, 0), "up": (0, 0, 1), "hfov": 40,
                             "res": (16, 16)},
                     params={"max_path_length": 4})
    o = O.OracleScene(sc)
    img, _ = o.render(0, 512, threads=8)
    img /= 512
    centre = img[5:11, 5:11]
    assert abs(centre.mean() - 0.5) < 0.01
    assert np.allclose(img[0, 0], 1.0)


@pytest.mark.parametrize("alg", ["ptn", "pt"])
def test_path_tracers_agree_on_box(alg):
    """The MIS path tracer (NEE + BSDF sampling) and the naive path tracer estimate the same
    integral: block means of the box agree within noise."""
    ref = _box_with("pt", res=32, maxpl=4)
    other = _box_with(alg, res=32, maxpl=4)
    a, _ = O.OracleScene(ref).render(0, 1024, threads=8)
    other.set_params(seed=99)
    b, _ = O.OracleScene(other).render(0, 2048 if alg == "ptn" else 1024, threads=8)
    a /= 1024
    b /= (2048 if alg == "ptn" else 1024)
    ba = a.reshape(4, 8, 4, 8, 3).mean(axis=(1, 3))
    bb = b.reshape(4, 8, 4, 8, 3).mean(axis=(1, 3))
    rel = np.abs(ba - bb) / np.maximum(ba, 1e-3)
    assert np.median(rel) < 0.03
    assert abs(a.mean() - b.mean()) / a.mean() < 0.02


def test_direct_illumination_estimators_agree():
    """dbs, dlss, dlsa and dmis estimate the same one-bounce integral (direct_illumination.hxx)."""
    means = {}
    for alg in ("dmis", "dlss", "dlsa", "dbs"):
        sb = _box_with(alg, res=32, maxpl=0)
        img, _ = O.OracleScene(sb).render(0, 2048 if alg == "dbs" else 512, threads=8)
        means[alg] = img.mean() / (2048 if alg == "dbs" else 512)
    for alg, m in means.items():
        assert abs(m - means["dmis"]) / means["dmis"] < 0.02, (alg, means)


# ---- scene intersection (scene.hxx:42-74, scene_graph.hxx:33-245) -----------------------------
from geom64 import cast as cast64, random_rays  # noqa: E402


@pytest.mark.parametrize("workload", [1, 2])
def test_oracle_intersect_against_fp64_ray_casting(workload):
    """The oracle's closest hit (float32, the reference's edge-sign test and FP64 sphere roots)
    against an independent float64 ray caster: identical hit/miss and material, distance and
    normal to float32 rounding, except on rays within 1e-5 of an edge, a tangent, a bound or a
    second surface."""
    sb = SceneBuild.workload(workload, seed=41, res=8)
    orc = O.OracleScene(sb)
    rng = np.random.default_rng(41 + workload)
    # half the rays start inside the box, half outside it (walls hit from both sides)
    org_in, dir_in = random_rays(rng, 10000)
    org_out, dir_out = random_rays(rng, 10000, -1.0, 2.0)
    org, dir = np.concatenate([org_in, org_out]), np.concatenate([dir_in, dir_out])
    tmin = rng.choice([0.0, 1e-3, 0.05], len(org)).astype(np.float32)
    dist, nrm, mat = orc.intersect(org, dir, tmin)
    t64, m64, n64, margin = cast64(sb, org, dir, tmin)
    clear = margin > 1e-5
    assert clear.mean() > 0.99
    hit64 = np.isfinite(t64)
    assert np.array_equal((dist >= 0)[clear], hit64[clear])
    assert np.array_equal(mat[clear], m64[clear])
    both = clear & hit64
    assert both.mean() > 0.3
    rel = np.abs(dist[both] - t64[both]) / np.maximum(t64[both], 1e-2)
    assert rel.max() < 2e-5, rel.max()
    assert np.abs(nrm[both] - n64[both]).max() < 2e-5
    # a plausible slip (a dropped sphere root or one-sided triangles) would show here
    assert (mat[clear & hit64] >= 0).all() and len(np.unique(m64[both])) >= 5


def test_oracle_occluded_against_fp64_ray_casting():
    """Occluded(p, dir, tmax) = a hit between p + 1e-3 dir and tmax - 2e-3 (scene.hxx:61-74)."""
    sb = SceneBuild.workload(2, seed=42, res=8)
    orc = O.OracleScene(sb)
    rng = np.random.default_rng(43)
    p_in, dir_in = random_rays(rng, 10000)
    p_out, dir_out = random_rays(rng, 10000, -1.0, 2.0)
    p, dir = np.concatenate([p_in, p_out]), np.concatenate([dir_in, dir_out])
    tmax = rng.uniform(0.05, 1.5, len(p)).astype(np.float32)
    occ = orc.occluded(p, dir, tmax)
    org = p.astype(np.float64) + 1e-3 * dir.astype(np.float64)
    t64, _, _, margin = cast64(sb, org, dir, 0.0, tmax.astype(np.float64) - 2e-3)
    clear = margin > 1e-5
    assert clear.mean() > 0.99
    assert np.array_equal(occ[clear], np.isfinite(t64)[clear])
    assert 0.1 < occ.mean() < 0.8
