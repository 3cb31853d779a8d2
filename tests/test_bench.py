"""The measurement contract of bench.py on the CPU: the one-GPU rule and the roofline model."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from pg3render_amd import SceneBuild  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_roofline_model_of_the_headline_scene():
    """The box + steerable-EM scene: 10 triangles and 2 spheres, so 2 x (10 x 62 + 2 x 35) + 400 =
    1,780 FLOP per segment; 112 B per sample + 304 B per segment; the bound is the larger fraction."""
    d = SceneBuild.workload(2, seed=2024).desc
    assert (d.num_triangles, d.num_spheres) == (10, 2)
    samples, segments, t = 16_777_216, 78_381_311, 10.4e-3
    bound, legs = bench.roofline_legs(d, samples, segments, t)
    assert legs["valu"]["flop_per_segment"] == 1780
    assert legs["hbm"]["bytes_per_launch"] == 112 * samples + 304 * segments
    flops = samples * 690 + segments * 1780
    assert legs["valu"]["flop_per_launch"] == flops
    assert legs["hbm"]["frac"] == pytest.approx((112 * samples + 304 * segments) / t / 1e9 / 8000.0, abs=1e-4)
    assert legs["valu"]["frac"] == pytest.approx(flops / t / 1e12 / 157.3, abs=1e-4)
    assert bound == max(legs, key=lambda k: legs[k]["frac"]) == "hbm"


def test_measured_constants_follow_the_calibration():
    """Traffic = FETCH_SIZE x 2 + WRITE_SIZE x 1 KiB (the factors of a 512 MiB copy); the issue
    peak is 2 wave instructions per CU per cycle on 256 CUs at 2.4 GHz."""
    assert bench.MEASURED_TRAFFIC[(2, "megakernel")] == int((2.0 * 3855186 + 1467896) * 1024)
    assert bench.VALU_ISSUE_PEAK == pytest.approx(1.2288e12)
    for name, traffic in bench.MEASURED_SPLIT_TRAFFIC.items():
        assert traffic > 0, name
