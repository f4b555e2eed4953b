"""The space partitioning inlier selector (inc/inlier_selectors/
space_partitioning_ransac.h:339-504) in the engine's homography loop, alone
and with SPRT (findHomography's defaults), against the oracle run with the
same seed: identical inlier masks and iteration / LO / graph-cut counts,
models within 1e-12.

The device selects each candidate's cells (in the order the reference's grid
container iterates them) and scores the kept cells' points in that order;
with SPRT the walk tests the kept points cell by cell.
"""
import numpy as np
import pytest

from pygcransac import synthetic

pytestmark = pytest.mark.gpu


def _same(st, so, M, Mo, mask, mo, rc):
    assert np.array_equal(mask, mo)
    for key in ("iteration_number", "local_optimization_number", "graph_cut_number"):
        assert st[key] == so[key], key
    if rc > 0:
        np.testing.assert_allclose(M, Mo, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("div,use_sprt,seed", [(4, False, 0), (8, False, 1), (4, True, 2), (8, True, 3)])
def test_space_partitioning_matches_oracle(oracle, div, use_sprt, seed):
    import pygcransac
    sc = synthetic.homography_scene(synthetic.scene_seed(17, seed), n=3000, inlier_ratio=[0.4, 0.25, 0.5, 0.35][seed])
    print("space partitioning div=%d sprt=%d seed=%d" % (div, use_sprt, seed), flush=True)
    H, mask = pygcransac.findHomography(sc["corr"], 480, 640, 480, 640, None, threshold=3.0, conf=0.99,
                                        spatial_coherence_weight=0.975, max_iters=10000, min_iters=50,
                                        use_sprt=use_sprt, sampler=0, neighborhood=0, neighborhood_size=float(div),
                                        use_space_partitioning=True, lo_number=50, seed=seed)
    st = pygcransac.last_stats()
    p = oracle.params(3.0, 0.99, 0.975, 10000, 50, 0, 50, float(div), seed, use_sprt=use_sprt,
                      use_space_partitioning=True)
    rc, Ho, mo, so = oracle.find_homography(sc["corr"], 480, 640, 480, 640, p)
    assert rc > 0 or use_sprt
    _same(st, so, H, Ho, mask, mo, rc)


def test_find_homography_reference_defaults_match_oracle(oracle):
    """Every keyword at the reference's default (PROSAC, SPRT, space
    partitioning, grid of 4 divisions) except the threshold and the seed."""
    import pygcransac
    sc = synthetic.homography_scene(synthetic.scene_seed(17, 10), n=2500, inlier_ratio=0.6)
    H, mask = pygcransac.findHomography(sc["corr"], 480, 640, 480, 640, None, threshold=3.0, seed=5)
    st = pygcransac.last_stats()
    p = oracle.params(3.0, 0.99, 0.975, 10000, 50, 1, 50, 4.0, 5, use_sprt=True, use_space_partitioning=True)
    rc, Ho, mo, so = oracle.find_homography(sc["corr"], 480, 640, 480, 640, p)
    _same(st, so, H, Ho, mask, mo, rc)


def test_batch_mixing_selector_and_plain_pairs(oracle):
    import pygcransac
    probs, refs = [], []
    for i in range(6):
        sc = synthetic.homography_scene(synthetic.scene_seed(17, 20 + i), n=2000, inlier_ratio=0.45)
        sp, sprt = i % 2 == 0, i % 3 == 0
        probs.append(dict(kind="homography", correspondences=sc["corr"], h1=480, w1=640, h2=480, w2=640,
                          threshold=3.0, conf=0.99, spatial_coherence_weight=0.975, max_iters=10000, min_iters=50,
                          use_sprt=sprt, sampler=0, neighborhood=0, neighborhood_size=6.0,
                          use_space_partitioning=sp, lo_number=50, seed=60 + i))
        p = oracle.params(3.0, 0.99, 0.975, 10000, 50, 0, 50, 6.0, 60 + i, use_sprt=sprt, use_space_partitioning=sp)
        refs.append(oracle.find_homography(sc["corr"], 480, 640, 480, 640, p))
    res = pygcransac.findBatch(probs)
    stats = pygcransac.last_batch_stats()
    for i, ((M, mask), (rc, Mo, mo, so)) in enumerate(zip(res, refs)):
        assert stats[i]["status"] == rc, i
        _same(stats[i], so, M, Mo, mask, mo, rc)
