"""GPU parity of the homography path against the CPU oracle.

Bar: bit-exact for counts, masks, sample statuses and labellings; FP64 values
and models equal to the oracle's bits (both sides evaluate the same IEEE
operations in the same order without fused multiply-add).
"""
import numpy as np
import pytest

from pygcransac import synthetic

pytestmark = pytest.mark.gpu

TAU2_H = (1.5 * 3.0) ** 2


def _bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def _scene(seed):
    return synthetic.homography_scene(synthetic.scene_seed(0, seed))


def _models_from_samples(oracle, pts, rng, k):
    samples = np.stack([rng.choice(len(pts), 4, replace=False) for _ in range(k)]).astype(np.int32)
    models, status = oracle.h4_solve(pts, samples)
    return samples, models, status


@pytest.mark.parametrize("seed", [0, 1])
def test_minimal_solver_bit_exact(oracle, engine, seed):
    sc = _scene(seed)
    pts = sc["corr"]
    rng = np.random.default_rng(seed)
    samples = np.stack([rng.choice(len(pts), 4, replace=False) for _ in range(3000)]).astype(np.int32)
    om, ost = oracle.h4_solve(pts, samples)
    dm, dst = engine.solve_minimal(0, pts, samples)
    assert np.array_equal(ost, dst)
    ok = ost == 2
    assert ok.sum() > 100
    assert np.array_equal(om[ok].view(np.uint64), dm[ok].view(np.uint64))


def test_minimal_solver_degenerate_sample(oracle, engine):
    # three collinear points: the oriented test or the elimination rejects it
    pts = np.array([[0, 0, 0, 0], [1, 1, 1, 1], [2, 2, 2, 2], [3, 3, 3, 3], [5, 1, 4, 2]], np.float64) * 10.0
    samples = np.array([[0, 1, 2, 3], [0, 1, 2, 4]], np.int32)
    om, ost = oracle.h4_solve(pts, samples)
    dm, dst = engine.solve_minimal(0, pts, samples)
    assert np.array_equal(ost, dst)
    assert (ost != 2).all()


@pytest.mark.parametrize("seed", [0, 3])
def test_scoring_bit_exact(oracle, engine, seed):
    sc = _scene(seed)
    pts = sc["corr"]
    rng = np.random.default_rng(100 + seed)
    _, models, status = _models_from_samples(oracle, pts, rng, 400)
    models = np.concatenate([models[status == 2], (sc["H"] / sc["H"][2, 2]).reshape(1, 9)])
    oc, ov, omask = oracle.score(0, pts, models, TAU2_H)
    dc, dv, dmask = engine.score_models(0, pts, models, TAU2_H, want_masks=True)
    assert np.array_equal(oc, dc)
    assert np.array_equal(omask, dmask)
    assert np.array_equal(ov.view(np.uint64), dv.view(np.uint64))
    assert dc[-1] > 900  # the true model finds the inliers


@pytest.mark.parametrize("seed", [0, 2])
def test_labeling_matches_oracle(oracle, engine, seed):
    sc = _scene(seed)
    pts = sc["corr"]
    H = sc["H"] / sc["H"][2, 2]
    eps = np.finfo(np.float64).eps
    cell = np.array([(640 + eps) / 8, (480 + eps) / 8, (640 + eps) / 8, (480 + eps) / 8])
    rng = np.random.default_rng(seed)
    for k in range(4):
        M = H * (1.0 + 0.002 * k * rng.standard_normal((3, 3)))
        M[2, 2] = 1.0
        os_ = oracle.labeling(0, pts, M, 0.975, 3.0, cell, 8)
        ds = engine.labeling(0, pts, M, 0.975, 3.0, cell, 8).astype(bool)
        assert np.array_equal(os_, ds), (k, np.nonzero(os_ != ds))


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_find_homography_matches_oracle(oracle, seed):
    import pygcransac
    sc = _scene(seed)
    corr = sc["corr"]
    p = oracle.params(3.0, 0.99, 0.975, 10000, 50, 0, 50, 8.0, seed)
    rc, Ho, mo, so = oracle.find_homography(corr, 480, 640, 480, 640, p)
    Hd, md = pygcransac.findHomography(corr, 480, 640, 480, 640, None, threshold=3.0, conf=0.99,
                                       spatial_coherence_weight=0.975, max_iters=10000, min_iters=50, use_sprt=False,
                                       sampler=0, neighborhood=0, neighborhood_size=8, use_space_partitioning=False,
                                       lo_number=50, seed=seed)
    sd = pygcransac.last_stats()
    assert rc == md.sum()
    assert np.array_equal(mo, md)
    assert so["iteration_number"] == sd["iteration_number"]
    assert so["local_optimization_number"] == sd["local_optimization_number"]
    assert so["graph_cut_number"] == sd["graph_cut_number"]
    np.testing.assert_allclose(Hd, Ho, rtol=1e-12, atol=1e-12)
    assert (md & sc["gt"]).sum() >= 0.99 * sc["gt"].sum()


def test_nonminimal_fit_beyond_lds_bit_exact(oracle, engine):
    # 4,800 inliers = 9,600 rows: the first rows of the system stay in LDS,
    # the rest in the fit's scratch (four-wave fit)
    sc = synthetic.homography_scene(synthetic.scene_seed(5, 3), n=8000, inlier_ratio=0.6)
    pts = sc["corr"]
    inl = np.nonzero(sc["gt"])[0]
    w = np.zeros(len(pts))
    w[inl] = np.random.default_rng(4).uniform(0, 1, len(inl))
    for sample, weights, wm in [(inl, None, 0), (inl, w, 1)]:
        ok, Ho = oracle.h_nonminimal(pts, sample, weights)
        dm = engine.fit_nonminimal(0, pts, sample.astype(np.int32), weights, wm, 1, 3.0)
        assert np.array_equal(_bits(Ho.ravel()), _bits(dm[0]))


def test_find_homography_large_scene_matches_oracle(oracle):
    import pygcransac
    sc = synthetic.homography_scene(synthetic.scene_seed(5, 4), n=8000, inlier_ratio=0.5)
    p = oracle.params(3.0, 0.99, 0.975, 10000, 50, 0, 50, 8.0, 5)
    rc, Ho, mo, so = oracle.find_homography(sc["corr"], 480, 640, 480, 640, p)
    Hd, md = pygcransac.findHomography(sc["corr"], 480, 640, 480, 640, None, threshold=3.0, conf=0.99,
                                       spatial_coherence_weight=0.975, max_iters=10000, min_iters=50, use_sprt=False,
                                       sampler=0, neighborhood=0, neighborhood_size=8, use_space_partitioning=False,
                                       lo_number=50, seed=5)
    sd = pygcransac.last_stats()
    assert rc == md.sum() and rc > 3500
    assert np.array_equal(mo, md)
    for key in ("iteration_number", "local_optimization_number", "graph_cut_number"):
        assert so[key] == sd[key], key
    np.testing.assert_allclose(Hd, Ho, rtol=1e-12, atol=1e-12)
