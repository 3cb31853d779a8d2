"""Host build of the steerable sampler tables (pg3g_steerable_build, no GPU).

The build decides each refinement level's subdivision tests on several threads and lists the
leaves in the reference's depth-first order (RefineEmTriangulation, em_steerable_sampler.hxx:
3090-3155).  Checks: the tables are identical for 1 and 8 threads, and equal to the oracle's
sequential restatement of the reference build on adaptive refinements (levels 2..6 and 4..7,
where the approximation-error test of :3219-3397 decides which triangles split).
"""
import numpy as np
import pytest

from pg3render_amd import SceneBuild, steerable_build
from oracle import oracle as O


def adaptive_scene(min_level, max_level, max_err=0.5):
    sb = SceneBuild.workload(3, seed=2024, variant=7)      # 256x128 sky + sun EM
    sb.desc.em.st_min_subdiv_level = min_level
    sb.desc.em.st_max_subdiv_level = max_level
    sb.desc.em.st_max_approx_error = max_err
    return sb


@pytest.mark.parametrize("levels", [(2, 6), (4, 7)])
def test_thread_count_does_not_change_the_tables(levels):
    sb = adaptive_scene(*levels)
    i1, l1, w1 = steerable_build(sb.desc.em, threads=1)
    i8, l8, w8 = steerable_build(sb.desc.em, threads=8)
    assert i1["threads"] == 1 and i8["threads"] == 8
    for k in ("num_leaves", "num_vertices", "num_levels"):
        assert i1[k] == i8[k]
    # adaptive: more leaves than the uniform min level, fewer than the uniform max level
    assert 20 * 4 ** levels[0] < i1["num_leaves"] < 20 * 4 ** levels[1]
    assert np.array_equal(l1, l8)
    assert np.array_equal(w1, w8)


@pytest.mark.parametrize("levels", [(2, 6), (4, 7)])
def test_matches_the_oracle_build(levels):
    sb = adaptive_scene(*levels)
    info, leaves, w = steerable_build(sb.desc.em, threads=4)
    leaves_o, nv_o, wo = O.OracleScene(sb).steerable()
    assert leaves.shape == leaves_o.shape
    assert info["num_vertices"] == nv_o
    assert np.max(np.abs(leaves - leaves_o)) <= 1e-6
    np.testing.assert_allclose(w, wo, rtol=1e-5, atol=1e-6 * np.abs(wo).max())


def test_uniform_levels_and_errors():
    sb = SceneBuild.workload(2, seed=5)                   # levels 4..4: 20 * 4^4 leaves
    info, leaves, _ = steerable_build(sb.desc.em, threads=2)
    assert info["num_leaves"] == 5120 and info["num_levels"] == 6
    assert np.allclose(np.linalg.norm(leaves, axis=2), 1.0, atol=1e-6)
    sb.desc.em.st_max_subdiv_level = 13
    sb.desc.em.st_min_subdiv_level = 13
    with pytest.raises(Exception, match="max subdivision level"):
        steerable_build(sb.desc.em, threads=2, leaves=False)
