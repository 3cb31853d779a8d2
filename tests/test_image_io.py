"""Image and EM files, image comparison and the steerable-table cache (host only, no GPU).

Follows Framebuffer::SaveHDR / SaveBMP (frame_buffer.hxx:114-200), EnvironmentMapImage::LoadImage
(em_image.hxx:51-93: azimuth rotation by int(FmodX(rot, 1) * width) texels and a scale), the
reference EM table of scene.hxx:850-943, and the .emssd cache it replaces
(em_steerable_sampler.hxx:1596-2030).  Expected values are computed independently here with
numpy (RGBE and PFM encoders written from the format specifications).
"""
import math
import os
import struct
import subprocess

import numpy as np
import pytest

from pg3render_amd import (SceneBuild, _ffi, compare_images, load_image, save_image, steerable_build,
                           steerable_cache_dir)
from pg3render_amd._ffi import Pg3gError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rgbe_encode(img):
    """Radiance RGBE of an (h, w, 3) image as SaveHDR computes it (truncating mantissas)."""
    v = img.max(axis=2).astype(np.float64)
    m, e = np.frexp(v)
    scale = np.where(v >= 1e-32, m * 256.0 / np.where(v > 0, v, 1), 0.0)
    out = np.zeros(img.shape[:2] + (4,), np.uint8)
    out[..., :3] = np.where(v[..., None] >= 1e-32, np.floor(img * scale[..., None]), 0).astype(np.uint8)
    out[..., 3] = np.where(v >= 1e-32, e + 128, 0).astype(np.uint8)
    return out


def rgbe_decode(b):
    e = b[..., 3].astype(np.int32)
    f = np.where(e > 0, np.ldexp(1.0, e - 136), 0.0)
    return ((b[..., :3].astype(np.float64) + 0.5) * f[..., None]).astype(np.float32)


def write_rle_hdr(path, img):
    """New-style run-length RGBE scanlines (runs of equal bytes as 128+n, literals as n)."""
    h, w, _ = img.shape
    enc = rgbe_encode(img)
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n")
        f.write(f"-Y {h} +X {w}\n".encode())
        for y in range(h):
            f.write(bytes([2, 2, w >> 8, w & 255]))
            for c in range(4):
                row = enc[y, :, c]
                x = 0
                while x < w:
                    run = 1
                    while x + run < w and run < 127 and row[x + run] == row[x]:
                        run += 1
                    if run >= 3:
                        f.write(bytes([128 + run, row[x]]))
                        x += run
                    else:
                        n = min(128, w - x)
                        lit = 1
                        while lit < n and not (x + lit + 2 < w and row[x + lit] == row[x + lit + 1] == row[x + lit + 2]):
                            lit += 1
                        f.write(bytes([lit]) + bytes(row[x:x + lit]))
                        x += lit


def write_pfm(path, img, little=True, grey=False):
    h, w = img.shape[:2]
    data = img[..., 0] if grey else img
    with open(path, "wb") as f:
        f.write(f"{'Pf' if grey else 'PF'}\n{w} {h}\n{-1.0 if little else 1.0}\n".encode())
        f.write(np.ascontiguousarray(data[::-1]).astype("<f4" if little else ">f4").tobytes())


def test_hdr_write_matches_rgbe_and_reads_back(tmp_path):
    rng = np.random.default_rng(1)
    img = (rng.random((7, 9, 3)) * rng.choice([1e-3, 1.0, 50.0], size=(7, 9, 1))).astype(np.float64)
    img[0, 0] = 0.0
    p = str(tmp_path / "a.hdr")
    save_image(p, img, scale=0.5)
    raw = open(p, "rb").read()
    header = b"#?RADIANCE\n# PG3Render\nFORMAT=32-bit_rle_rgbe\n\n-Y 7 +X 9\n"
    assert raw.startswith(header)
    body = np.frombuffer(raw[len(header):], np.uint8).reshape(7, 9, 4)
    assert np.array_equal(body, rgbe_encode(img * 0.5))
    back = load_image(p)
    assert np.array_equal(back, rgbe_decode(body))
    assert np.all(np.abs(back - img * 0.5) <= (img * 0.5).max(axis=2, keepdims=True) / 256 + 1e-30)


def test_run_length_hdr_reads_like_flat(tmp_path):
    rng = np.random.default_rng(2)
    img = rng.random((5, 40, 3)).astype(np.float32) * 3
    img[:, 10:30] = img[:, 10:11]                  # long runs
    p = str(tmp_path / "rle.hdr")
    write_rle_hdr(p, img)
    assert np.array_equal(load_image(p), rgbe_decode(rgbe_encode(img)))


@pytest.mark.parametrize("little,grey", [(True, False), (False, False), (True, True)])
def test_pfm(tmp_path, little, grey):
    rng = np.random.default_rng(3)
    img = rng.standard_normal((6, 11, 3)).astype(np.float32)
    p = str(tmp_path / "a.pfm")
    write_pfm(p, img, little, grey)
    want = np.repeat(img[..., :1], 3, axis=2) if grey else img
    assert np.array_equal(load_image(p), want)


def test_raw_and_em_rotation(tmp_path):
    rng = np.random.default_rng(4)
    img = rng.random((4, 10, 3)).astype(np.float32)
    p = str(tmp_path / "a.raw")
    with open(p, "wb") as f:
        f.write(struct.pack("<II", 10, 4) + img.astype("<f4").tobytes())
    assert np.array_equal(load_image(p), img)
    for rot, scale in [(0.0, 1.0), (0.25, 2.0), (-0.12, 0.5), (1.83, 1.5)]:
        i_rot = int(math.fmod(rot, 1.0) % 1.0 * 10) if math.fmod(rot, 1.0) >= 0 else int((math.fmod(rot, 1.0) + 1.0) * 10)
        want = np.roll(img, i_rot, axis=1) * np.float32(scale)
        assert np.array_equal(load_image(p, em_rotation=rot, em_scale=scale), want), rot


def test_bmp(tmp_path):
    rng = np.random.default_rng(5)
    img = rng.random((3, 5, 3)) * 1.5                # width 5: rows padded to 16 bytes
    p = str(tmp_path / "a.bmp")
    save_image(p, img, scale=1.0, gamma=2.2)
    raw = open(p, "rb").read()
    assert raw[:2] == b"BM"
    size, _, off, hsz, w, h, planes, bpp = struct.unpack("<IIIIiiHH", raw[2:30])
    assert (off, hsz, w, h, planes, bpp) == (54, 40, 5, 3, 1, 24) and size == len(raw) == 54 + 16 * 3
    for y in range(3):
        row = np.frombuffer(raw[54 + 16 * y: 54 + 16 * y + 15], np.uint8).reshape(5, 3)
        src = img[3 - 1 - y]                          # bottom row first
        want = np.clip(np.power(src, 1 / 2.2) * 255.0, 0, 255).astype(np.uint8)[:, ::-1]   # BGR
        assert np.array_equal(row, want)


def test_compare_images_closed_forms():
    rng = np.random.default_rng(6)
    ref = (rng.random((16, 16, 3)) + 0.1).astype(np.float32)
    d, diff = compare_images(ref, ref)
    assert d["pixel_rel_rmse"] == 0 and d["image_rel_rmse"] == 0 and not diff.any()
    d, diff = compare_images(ref * np.float32(1.02), ref)
    assert abs(d["pixel_rel_rmse"] - 0.02) < 1e-5 and abs(d["image_rel_rmse"] - 0.02) < 1e-5
    assert np.allclose(diff, ref * 0.02, rtol=1e-4)


def test_pg3diff_exit_status(tmp_path):
    exe = os.path.join(ROOT, "pg3render_amd", "pg3diff")
    if not os.path.exists(exe):
        pytest.skip("pg3diff not built")
    rng = np.random.default_rng(7)
    ref = rng.random((8, 8, 3)) + 0.5
    a, b, c = (str(tmp_path / n) for n in ("ref.hdr", "close.hdr", "far.hdr"))
    save_image(a, ref)
    save_image(b, ref * 1.01)
    save_image(c, ref * 1.5)
    assert subprocess.run([exe, "-threshold", "0.05", b, a], capture_output=True).returncode == 0
    r = subprocess.run([exe, "-threshold", "0.05", "-output", str(tmp_path / "d.hdr"), c, a], capture_output=True)
    assert r.returncode == 1 and b"FAILED" in r.stdout and os.path.exists(tmp_path / "d.hdr")


def test_reference_scenes_with_image_ems(tmp_path):
    for em, (w, h) in [(2, (8, 4)), (3, (1024, 512))]:
        sb = SceneBuild.reference(22, em_type=em)
        t = sb.em_texels()
        assert t.shape == (h, w, 3) and np.all(t == 1.0)
        assert sb.desc.em.sampler == _ffi.EM_SIMPLE and sb.desc.em.bilinear == 1
    with pytest.raises(Pg3gError):
        SceneBuild.reference(22, em_type=5)          # a photograph: the caller names the file
    img = np.random.default_rng(8).random((16, 32, 3)).astype(np.float32)
    p = str(tmp_path / "probe.pfm")
    write_pfm(p, img)
    sb = SceneBuild.reference(22, em_type=7, em_path=p)   # kEMImgDoge2: rotation 0.83, scale 1.5
    assert np.array_equal(sb.em_texels(), load_image(p, em_rotation=0.83, em_scale=1.5))


def test_steerable_cache(tmp_path):
    sb = SceneBuild.workload(3, seed=2024, variant=7)
    em = sb.desc.em
    em.st_min_subdiv_level, em.st_max_subdiv_level, em.st_max_approx_error = 2, 6, 0.5
    try:
        steerable_cache_dir(str(tmp_path))
        i0, l0, w0 = steerable_build(em, threads=4)
        files = sorted(os.listdir(tmp_path))
        assert not i0["from_cache"] and len(files) == 1 and files[0].endswith(".pg3st")
        i1, l1, w1 = steerable_build(em, threads=4)
        assert i1["from_cache"] and np.array_equal(l0, l1) and np.array_equal(w0, w1)
        assert i1["num_vertices"] == i0["num_vertices"] and i1["num_levels"] == i0["num_levels"]
        # other content, same name-independent parameters: a new entry
        sb.em_texels()[0, 0, 0] += 1.0
        i2, _, _ = steerable_build(em, threads=4, leaves=False)
        assert not i2["from_cache"] and len(os.listdir(tmp_path)) == 2
        sb.em_texels()[0, 0, 0] -= 1.0
        # a damaged entry is rebuilt and replaced
        path = os.path.join(tmp_path, files[0])
        data = open(path, "rb").read()
        open(path, "wb").write(data[:-9] + bytes([data[-9] ^ 1]) + data[-8:])
        i3, l3, _ = steerable_build(em, threads=4)
        assert not i3["from_cache"] and np.array_equal(l3, l0)
        assert steerable_build(em, threads=4, leaves=False)[0]["from_cache"]
    finally:
        steerable_cache_dir(None)
    assert not steerable_build(em, threads=4, leaves=False)[0]["from_cache"]
