"""GPU parity of the probability-consuming samplers: importance (3,
inc/samplers/importance_sampler.h:51-135) and adaptive reordering (4,
inc/samplers/adaptive_reordering_sampler.h:46-166), through the single-pair
entry points, against the oracle run with the same seed and probabilities.
Bar: identical inlier masks and iteration / LO / graph-cut counts, models
within 1e-12."""
import numpy as np
import pytest

from pygcransac import synthetic

pytestmark = pytest.mark.gpu


def _probs(gt, seed, informative=True):
    rng = np.random.default_rng(seed)
    if informative:
        return np.where(gt, rng.uniform(0.4, 1.0, len(gt)), rng.uniform(0.0, 0.7, len(gt)))
    return rng.uniform(0.0, 1.0, len(gt))


@pytest.mark.parametrize("sampler", [3, 4])
@pytest.mark.parametrize("seed", [0, 1])
def test_homography_probability_samplers_match_oracle(oracle, sampler, seed):
    import pygcransac
    sc = synthetic.homography_scene(synthetic.scene_seed(6, seed), n=1800, inlier_ratio=0.35)
    pr = _probs(sc["gt"], seed, informative=seed == 0)
    print("homography sampler=%d seed=%d" % (sampler, seed), flush=True)
    H, mask = pygcransac.findHomography(sc["corr"], 480, 640, 480, 640, pr, threshold=3.0, conf=0.99,
                                        spatial_coherence_weight=0.975, max_iters=5000, min_iters=50,
                                        use_sprt=False, sampler=sampler, neighborhood=0, neighborhood_size=8,
                                        use_space_partitioning=False, lo_number=50, sampler_variance=0.1,
                                        seed=seed)
    st = pygcransac.last_stats()
    p = oracle.params(3.0, 0.99, 0.975, 5000, 50, sampler, 50, 8.0, seed, probabilities=pr, sampler_variance=0.1)
    rc, Ho, mo, so = oracle.find_homography(sc["corr"], 480, 640, 480, 640, p)
    assert np.array_equal(mask, mo)
    for key in ("iteration_number", "local_optimization_number", "graph_cut_number"):
        assert st[key] == so[key], key
    if rc > 0:
        np.testing.assert_allclose(H, Ho, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("sampler", [3, 4])
def test_fundamental_probability_samplers_match_oracle(oracle, sampler):
    import pygcransac
    sc = synthetic.fundamental_scene(synthetic.scene_seed(6, 40 + sampler), n=2500, inlier_ratio=0.4)
    pr = _probs(sc["gt"], sampler)
    print("fundamental sampler=%d" % sampler, flush=True)
    F, mask = pygcransac.findFundamentalMatrix(sc["corr"], 480, 640, 480, 640, pr, threshold=1.0, conf=0.99,
                                               spatial_coherence_weight=0.975, max_iters=5000, min_iters=50,
                                               use_sprt=False, sampler=sampler, neighborhood=0,
                                               neighborhood_size=8, lo_number=50, sampler_variance=0.1, seed=9)
    st = pygcransac.last_stats()
    p = oracle.params(1.0, 0.99, 0.975, 5000, 50, sampler, 50, 8.0, 9, probabilities=pr, sampler_variance=0.1)
    rc, Fo, mo, so = oracle.find_fundamental(sc["corr"], 480, 640, 480, 640, p)
    assert np.array_equal(mask, mo)
    for key in ("iteration_number", "local_optimization_number", "graph_cut_number"):
        assert st[key] == so[key], key
    if rc > 0:
        np.testing.assert_allclose(F, Fo, rtol=1e-12, atol=1e-14)
    assert (mask & sc["gt"]).sum() >= 0.9 * sc["gt"].sum()


@pytest.mark.parametrize("sampler", [3, 4])
def test_essential_probability_samplers_match_oracle(oracle, sampler):
    import pygcransac
    sc = synthetic.essential_scene(synthetic.scene_seed(6, 60 + sampler), n=2000, inlier_ratio=0.4)
    pr = _probs(sc["gt"], 10 + sampler)
    print("essential sampler=%d" % sampler, flush=True)
    E, mask = pygcransac.findEssentialMatrix(sc["corr"], sc["K1"], sc["K2"], 768, 1024, 768, 1024, pr,
                                             threshold=0.75, conf=0.99, spatial_coherence_weight=0.975,
                                             max_iters=5000, min_iters=50, use_sprt=False, sampler=sampler,
                                             neighborhood=0, neighborhood_size=8, lo_number=50,
                                             sampler_variance=0.1, seed=4)
    st = pygcransac.last_stats()
    p = oracle.params(0.75, 0.99, 0.975, 5000, 50, sampler, 50, 8.0, 4, probabilities=pr, sampler_variance=0.1)
    rc, Eo, mo, so = oracle.find_essential(sc["corr"], sc["K1"], sc["K2"], 768, 1024, 768, 1024, p)
    assert np.array_equal(mask, mo)
    for key in ("iteration_number", "local_optimization_number", "graph_cut_number"):
        assert st[key] == so[key], key
    if rc > 0:
        np.testing.assert_allclose(E, Eo, rtol=1e-12, atol=1e-14)
