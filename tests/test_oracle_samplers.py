"""The probability-consuming samplers (importance, inc/samplers/
importance_sampler.h:51-135; adaptive reordering, inc/samplers/
adaptive_reordering_sampler.h:46-166) in the oracle, and the "no model"
answer of the single-pair entry points when their probabilities do not match
(py/gcransac_python.cpp:4369-4379).  CPU only."""
import ctypes

import numpy as np
import pytest

from pygcransac import _native, synthetic


def _h_scene(seed, n=1500, eps=0.3):
    sc = synthetic.homography_scene(synthetic.scene_seed(5, seed), n=n, inlier_ratio=eps)
    # informative probabilities: generator inliers high, outliers low, noisy
    rng = np.random.default_rng(seed)
    pr = np.where(sc["gt"], rng.uniform(0.5, 1.0, n), rng.uniform(0.0, 0.6, n))
    return sc, pr


@pytest.mark.parametrize("sampler", [3, 4])
def test_probability_samplers_recover_the_homography(oracle, sampler):
    """With informative probabilities both samplers find the plane: the
    returned inliers keep >= 95 % of the generator's inliers."""
    for seed in range(3):
        sc, pr = _h_scene(seed)
        p = oracle.params(3.0, 0.99, 0.975, 5000, 50, sampler, 50, 8.0, seed, probabilities=pr,
                          sampler_variance=0.1)
        rc, H, mask, st = oracle.find_homography(sc["corr"], 480, 640, 480, 640, p)
        assert rc > 0
        assert (mask & sc["gt"]).sum() >= 0.95 * sc["gt"].sum()


def test_adaptive_reordering_samples_the_most_probable_points_first(oracle):
    """The first sample of the adaptive-reordering sampler is the m points of
    highest probability: on a scene whose inliers carry the top
    probabilities, a run capped at one iteration already returns the plane,
    where the uniform sampler's single sample (all four inliers with
    probability 0.3^4) does not."""
    sc, _ = _h_scene(7)
    pr = np.where(sc["gt"], 0.9, 0.1) + np.random.default_rng(1).uniform(0, 1e-3, len(sc["gt"]))
    p = oracle.params(3.0, 0.99, 0.975, 1, 1, 4, 50, 8.0, 3, probabilities=pr, sampler_variance=0.1)
    rc, H, mask, st = oracle.find_homography(sc["corr"], 480, 640, 480, 640, p)
    assert rc > 0 and st["iteration_number"] == 1
    assert (mask & sc["gt"]).sum() >= 0.95 * sc["gt"].sum()
    pu = oracle.params(3.0, 0.99, 0.975, 1, 1, 0, 50, 8.0, 3)
    rcu, _, masku, _ = oracle.find_homography(sc["corr"], 480, 640, 480, 640, pu)
    assert (masku & sc["gt"]).sum() < 0.5 * sc["gt"].sum()


def test_importance_sampler_follows_the_probabilities(oracle):
    """Points of probability 0 are never drawn: with every outlier at 0, a run
    capped at one iteration returns the plane."""
    sc, _ = _h_scene(11, eps=0.2)
    pr = np.where(sc["gt"], 1.0, 0.0)
    p = oracle.params(3.0, 0.99, 0.975, 1, 1, 3, 50, 8.0, 5, probabilities=pr)
    rc, H, mask, st = oracle.find_homography(sc["corr"], 480, 640, 480, 640, p)
    assert rc > 0
    assert (mask & sc["gt"]).sum() >= 0.95 * sc["gt"].sum()


def test_probability_samplers_without_probabilities_return_no_model():
    """The reference drivers return 0 inliers when samplers 3 / 4 get no (or
    the wrong number of) probabilities; so does the engine, before any device
    work."""
    import pygcransac
    sc, pr = _h_scene(1, n=200)
    for sampler in (3, 4):
        H, mask = pygcransac.findHomography(sc["corr"], 480, 640, 480, 640, None, threshold=3.0, sampler=sampler,
                                            use_sprt=False, neighborhood=0, seed=1)
        assert H is None and mask.shape == (200,) and not mask.any()
        H, mask = pygcransac.findHomography(sc["corr"], 480, 640, 480, 640, pr[:100], threshold=3.0,
                                            sampler=sampler, use_sprt=False, neighborhood=0, seed=1)
        assert H is None and not mask.any()


def test_negative_probabilities_are_refused():
    import pygcransac
    sc, pr = _h_scene(2, n=200)
    pr = pr.copy()
    pr[5] = -0.5
    with pytest.raises(ValueError):
        try:
            pygcransac.findHomography(sc["corr"], 480, 640, 480, 640, pr, threshold=3.0, sampler=3,
                                      use_sprt=False, neighborhood=0, seed=1)
        except RuntimeError as e:  # no device: the argument check must come first
            raise AssertionError(str(e))
