"""The radius neighbourhood (inc/neighborhood/flann_neighborhood_graph.h:
100-139, searched exhaustively) and its min-cut on the device, against the
oracle run with the same seed: identical inlier masks and iteration / LO /
graph-cut counts, models within 1e-12.  Also grid cells beyond the LDS
min-cut's 2,048 points, which run with their node state in global memory.

Cases: homography with space partitioning off (findHomography_ builds the
radius graph only then), fundamental with a positive spatial weight,
essential with the reference's defaults (spatial weight 0.1, radius 20 px
on the pixel coordinates), a dense graph whose components hold hundreds of
points, a radius without any edge (no model, as the drivers return), and a
batch mixing grid and radius pairs.
"""
import numpy as np
import pytest

from pygcransac import synthetic

pytestmark = pytest.mark.gpu


def _same(st, so, M, Mo, mask, mo, rc, atol=1e-12):
    assert np.array_equal(mask, mo)
    for key in ("iteration_number", "local_optimization_number", "graph_cut_number"):
        assert st[key] == so[key], key
    if rc > 0:
        np.testing.assert_allclose(M, Mo, rtol=1e-12, atol=atol)


@pytest.mark.parametrize("radius,seed", [(20.0, 0), (45.0, 1)])
def test_homography_radius_graph_matches_oracle(oracle, radius, seed):
    import pygcransac
    sc = synthetic.homography_scene(synthetic.scene_seed(13, seed), n=2500, inlier_ratio=0.4)
    print("homography radius=%g seed=%d" % (radius, seed), flush=True)
    H, mask = pygcransac.findHomography(sc["corr"], 480, 640, 480, 640, None, threshold=3.0, conf=0.99,
                                        spatial_coherence_weight=0.975, max_iters=5000, min_iters=50,
                                        use_sprt=False, sampler=0, neighborhood=1, neighborhood_size=radius,
                                        use_space_partitioning=False, lo_number=50, seed=seed)
    st = pygcransac.last_stats()
    p = oracle.params(3.0, 0.99, 0.975, 5000, 50, 0, 50, radius, seed, neighborhood=1)
    rc, Ho, mo, so = oracle.find_homography(sc["corr"], 480, 640, 480, 640, p)
    assert rc > 0 and so["graph_cut_number"] > 0
    _same(st, so, H, Ho, mask, mo, rc)


def test_fundamental_radius_graph_matches_oracle(oracle):
    import pygcransac
    sc = synthetic.fundamental_scene(synthetic.scene_seed(13, 20), n=3000, inlier_ratio=0.35)
    F, mask = pygcransac.findFundamentalMatrix(sc["corr"], 480, 640, 480, 640, None, threshold=1.0, conf=0.99,
                                               spatial_coherence_weight=0.3, max_iters=5000, min_iters=50,
                                               use_sprt=False, sampler=0, neighborhood=1, neighborhood_size=25.0,
                                               lo_number=50, seed=5)
    st = pygcransac.last_stats()
    p = oracle.params(1.0, 0.99, 0.3, 5000, 50, 0, 50, 25.0, 5, neighborhood=1)
    rc, Fo, mo, so = oracle.find_fundamental(sc["corr"], 480, 640, 480, 640, p)
    assert rc > 0
    _same(st, so, F, Fo, mask, mo, rc, atol=1e-14)


@pytest.mark.parametrize("sampler,n,eps", [(0, 2500, 0.4), (1, 10000, 0.1)])
def test_essential_reference_defaults_match_oracle(oracle, sampler, n, eps):
    """findEssentialMatrix's neighbourhood defaults (spatial weight 0.1,
    radius graph of radius 20 px), including the 10,000-point / 10 %-inlier
    configuration."""
    import pygcransac
    sc = synthetic.essential_scene(synthetic.scene_seed(13, 40 + sampler), n=n, inlier_ratio=eps)
    print("essential sampler=%d n=%d" % (sampler, n), flush=True)
    E, mask = pygcransac.findEssentialMatrix(sc["corr"], sc["K1"], sc["K2"], 768, 1024, 768, 1024, None,
                                             threshold=0.75, conf=0.99, spatial_coherence_weight=0.1,
                                             max_iters=20000, min_iters=50, use_sprt=False, sampler=sampler,
                                             neighborhood=1, neighborhood_size=20.0, lo_number=50, seed=11)
    st = pygcransac.last_stats()
    p = oracle.params(0.75, 0.99, 0.1, 20000, 50, sampler, 50, 20.0, 11, neighborhood=1)
    rc, Eo, mo, so = oracle.find_essential(sc["corr"], sc["K1"], sc["K2"], 768, 1024, 768, 1024, p)
    assert rc > 0
    _same(st, so, E, Eo, mask, mo, rc, atol=1e-14)


def test_dense_radius_graph_matches_oracle(oracle):
    """1,500 correspondences squeezed into a 160 x 120 image pair with a
    25 px radius: about 25 neighbours per point."""
    import pygcransac
    sc = synthetic.homography_scene(synthetic.scene_seed(13, 60), n=1500, inlier_ratio=0.5)
    corr = sc["corr"] * 0.25
    off, nbr = oracle.radius_neighbors(corr, 25.0)
    assert len(nbr) > 20 * len(corr)
    H, mask = pygcransac.findHomography(corr, 120, 160, 120, 160, None, threshold=0.75, conf=0.99,
                                        spatial_coherence_weight=0.5, max_iters=5000, min_iters=50,
                                        use_sprt=False, sampler=0, neighborhood=1, neighborhood_size=25.0,
                                        use_space_partitioning=False, lo_number=50, seed=2)
    st = pygcransac.last_stats()
    p = oracle.params(0.75, 0.99, 0.5, 5000, 50, 0, 50, 25.0, 2, neighborhood=1)
    rc, Ho, mo, so = oracle.find_homography(corr, 120, 160, 120, 160, p)
    assert rc > 0 and so["graph_cut_number"] > 0
    _same(st, so, H, Ho, mask, mo, rc)


def test_radius_without_edges_returns_no_model(oracle):
    import pygcransac
    sc = synthetic.essential_scene(synthetic.scene_seed(13, 70), n=800, inlier_ratio=0.5)
    E, mask = pygcransac.findEssentialMatrix(sc["corr"], sc["K1"], sc["K2"], 768, 1024, 768, 1024, None,
                                             threshold=0.75, spatial_coherence_weight=0.1, neighborhood=1,
                                             neighborhood_size=1e-3, sampler=0, use_sprt=False, seed=1)
    assert E is None and mask.shape == (800,) and not mask.any()
    st = pygcransac.last_stats()
    assert st["iteration_number"] == 0
    p = oracle.params(0.75, 0.99, 0.1, 10000, 50, 0, 50, 1e-3, 1, neighborhood=1)
    rc, _, mo, so = oracle.find_essential(sc["corr"], sc["K1"], sc["K2"], 768, 1024, 768, 1024, p)
    assert rc == 0 and not mo.any()


def test_batch_mixing_grid_and_radius_pairs(oracle):
    import pygcransac
    probs, refs = [], []
    for i in range(6):
        seed = synthetic.scene_seed(13, 80 + i)
        radius = i % 2 == 0
        nb = dict(neighborhood=1, neighborhood_size=20.0) if radius else dict(neighborhood=0,
                                                                              neighborhood_size=8.0)
        if i % 3 == 2:
            sc = synthetic.essential_scene(seed, n=2000, inlier_ratio=0.3)
            pb = dict(kind="essential", correspondences=sc["corr"], K1=sc["K1"], K2=sc["K2"], h1=768, w1=1024,
                      h2=768, w2=1024, threshold=0.75, spatial_coherence_weight=0.1)
            p = oracle.params(0.75, 0.99, 0.1, 10000, 50, 0, 50, nb["neighborhood_size"], 30 + i,
                              neighborhood=nb["neighborhood"])
            refs.append(oracle.find_essential(sc["corr"], sc["K1"], sc["K2"], 768, 1024, 768, 1024, p))
        else:
            sc = synthetic.homography_scene(seed, n=1800, inlier_ratio=0.45)
            pb = dict(kind="homography", correspondences=sc["corr"], h1=480, w1=640, h2=480, w2=640,
                      threshold=3.0, spatial_coherence_weight=0.975, use_space_partitioning=False)
            p = oracle.params(3.0, 0.99, 0.975, 10000, 50, 0, 50, nb["neighborhood_size"], 30 + i,
                              neighborhood=nb["neighborhood"])
            refs.append(oracle.find_homography(sc["corr"], 480, 640, 480, 640, p))
        pb.update(nb, conf=0.99, max_iters=10000, min_iters=50, use_sprt=False, sampler=0, lo_number=50,
                  seed=30 + i)
        probs.append(pb)
    res = pygcransac.findBatch(probs)
    stats = pygcransac.last_batch_stats()
    for i, ((M, mask), (rc, Mo, mo, so)) in enumerate(zip(res, refs)):
        assert stats[i]["status"] == rc, i
        _same(stats[i], so, M, Mo, mask, mo, rc, atol=1e-14)


def test_grid_cell_beyond_lds_capacity_matches_oracle(oracle):
    """2,600 correspondences in one grid cell (one division per axis): the
    cell's min-cut runs with its node state in global memory."""
    import pygcransac
    sc = synthetic.homography_scene(synthetic.scene_seed(13, 90), n=2600, inlier_ratio=0.5)
    H, mask = pygcransac.findHomography(sc["corr"], 480, 640, 480, 640, None, threshold=3.0, conf=0.99,
                                        spatial_coherence_weight=0.975, max_iters=3000, min_iters=50,
                                        use_sprt=False, sampler=0, neighborhood=0, neighborhood_size=1.0,
                                        use_space_partitioning=False, lo_number=20, seed=4)
    st = pygcransac.last_stats()
    p = oracle.params(3.0, 0.99, 0.975, 3000, 50, 0, 20, 1.0, 4)
    rc, Ho, mo, so = oracle.find_homography(sc["corr"], 480, 640, 480, 640, p)
    assert rc > 0 and so["graph_cut_number"] > 0
    _same(st, so, H, Ho, mask, mo, rc)
