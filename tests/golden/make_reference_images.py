"""Generates the high-spp FP64 reference images used by the convergence tests.

The CPU oracle renders each workload at a small resolution in K independent batches (disjoint
iteration ranges of the counter-based RNG); the fixture stores the per-pixel mean radiance and
its standard error.  Run from the repository root after building (about 3 minutes on 8 cores):

    python tests/golden/make_reference_images.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from pg3render_amd import SceneBuild  # noqa: E402

CASES = [
    # name, workload, resolution, samples per pixel, batches, seed
    ("workload1_32px", 1, 32, 1 << 18, 16, 777),
    ("workload2_24px", 2, 24, 1 << 16, 16, 778),
]


def main():
    out_dir = os.path.dirname(os.path.abspath(__file__))
    for name, wl, res, spp, batches, seed in CASES:
        sb = SceneBuild.workload(wl, seed=2024, res=res)
        sb.set_params(seed=seed)
        orc = O.OracleScene(sb)
        per = spp // batches
        means = []
        t0 = time.time()
        for k in range(batches):
            img, _ = orc.render(k * per, per, threads=os.cpu_count())
            means.append(img / per)
        means = np.stack(means)
        mean = means.mean(axis=0)
        se = means.std(axis=0, ddof=1) / np.sqrt(batches)
        np.savez_compressed(os.path.join(out_dir, f"{name}.npz"), mean=mean, se=se, spp=spp, batches=batches,
                            seed=seed, scene_seed=2024, workload=wl, res=res)
        rel = se.sum(-1) / np.maximum(mean.sum(-1), 1e-9)
        print(f"{name}: {res}x{res} x {spp} spp in {time.time() - t0:.1f} s, median per-pixel relative SE "
              f"{np.median(rel):.4f}, max {rel.max():.4f}")


if __name__ == "__main__":
    main()
