"""GPU parity of the fundamental-matrix path against the CPU oracle.

Bar: bit-exact statuses, counts and masks; FP64 models and values with the
same bits as the oracle (identical IEEE operation sequences, no fused
multiply-add, identical reduction orders), and full runs with identical
inlier masks and statistics.
"""
import numpy as np
import pytest

from pygcransac import synthetic

pytestmark = pytest.mark.gpu

TAU2_F = (1.5 * 1.0) ** 2


def _scene(seed, n=2000):
    return synthetic.fundamental_scene(synthetic.scene_seed(1, seed), n=n, inlier_ratio=0.3)


def _bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


@pytest.mark.parametrize("seed", [0, 1])
def test_seven_point_bit_exact(oracle, engine, seed):
    pts = _scene(seed)["corr"]
    rng = np.random.default_rng(seed)
    samples = np.stack([rng.choice(len(pts), 7, replace=False) for _ in range(2000)]).astype(np.int32)
    om, ost = oracle.f7_solve(pts, samples)
    dm, dst = engine.solve_minimal(1, pts, samples)
    assert np.array_equal(ost, dst)
    ok = ost == 2
    assert ok.sum() > 200
    assert np.array_equal(_bits(om[ok]), _bits(dm[ok]))


def test_sampson_scoring_bit_exact(oracle, engine):
    sc = _scene(2)
    pts = sc["corr"]
    rng = np.random.default_rng(5)
    samples = np.stack([rng.choice(len(pts), 7, replace=False) for _ in range(200)]).astype(np.int32)
    om, ost = oracle.f7_solve(pts, samples)
    models = np.concatenate([om[ost == 2], sc["F"].reshape(1, 9)])
    oc, ov, omask = oracle.score(1, pts, models, TAU2_F)
    dc, dv, dmask = engine.score_models(1, pts, models, TAU2_F, want_masks=True)
    assert np.array_equal(oc, dc)
    assert np.array_equal(omask, dmask)
    assert np.array_equal(_bits(ov), _bits(dv))


@pytest.mark.parametrize("mode", ["lo", "irls", "post"])
def test_nonminimal_fit_bit_exact(oracle, engine, mode):
    sc = _scene(3)
    pts = sc["corr"]
    inl = np.nonzero(sc["gt"])[0]
    rng = np.random.default_rng(11)
    if mode == "lo":
        cases = [(rng.choice(inl, 49, replace=False), None, 0, 1) for _ in range(6)]
    elif mode == "irls":
        w = np.zeros(len(pts))
        w[inl] = rng.uniform(0.0, 1.0, len(inl)) ** 2
        cases = [(inl[:400], w, 1, 1), (inl, w, 1, 1)]
    else:
        F = sc["F"] / np.linalg.norm(sc["F"])
        cases = [(inl, F.ravel(), 2, 0)]
    for sample, w, wm, norm in cases:
        om = oracle.fit_nonminimal(1, pts, sample, w, wm, norm, 1.0)
        dm = engine.fit_nonminimal(1, pts, sample.astype(np.int32), w, wm, norm, 1.0)
        assert om.shape == dm.shape
        assert np.array_equal(_bits(om), _bits(dm)), np.abs(om - dm).max()


def test_homography_nonminimal_fit_bit_exact(oracle, engine):
    sc = synthetic.homography_scene(synthetic.scene_seed(0, 4))
    pts = sc["corr"]
    inl = np.nonzero(sc["gt"])[0]
    w = np.zeros(len(pts))
    w[inl] = np.random.default_rng(2).uniform(0, 1, len(inl))
    for sample, weights, wm in [(inl[:28], None, 0), (inl, None, 0), (inl, w, 1)]:
        ok, Ho = oracle.h_nonminimal(pts, sample, weights)
        dm = engine.fit_nonminimal(0, pts, sample.astype(np.int32), weights, wm, 1, 3.0)
        assert np.array_equal(_bits(Ho.ravel()), _bits(dm[0]))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_find_fundamental_matches_oracle(oracle, seed):
    import pygcransac
    sc = _scene(seed)
    corr = sc["corr"]
    p = oracle.params(1.0, 0.99, 0.975, 100000, 50, 0, 50, 8.0, seed)
    rc, Fo, mo, so = oracle.find_fundamental(corr, 480, 640, 480, 640, p)
    Fd, md = pygcransac.findFundamentalMatrix(corr, 480, 640, 480, 640, None, threshold=1.0, conf=0.99,
                                              spatial_coherence_weight=0.975, max_iters=100000, min_iters=50,
                                              use_sprt=False, sampler=0, neighborhood=0, neighborhood_size=8,
                                              lo_number=50, seed=seed)
    sd = pygcransac.last_stats()
    assert rc == md.sum()
    assert np.array_equal(mo, md)
    assert so["iteration_number"] == sd["iteration_number"]
    assert so["local_optimization_number"] == sd["local_optimization_number"]
    assert so["graph_cut_number"] == sd["graph_cut_number"]
    np.testing.assert_allclose(Fd, Fo, rtol=1e-12, atol=1e-14)
    assert (md & sc["gt"]).sum() >= 0.97 * sc["gt"].sum()


# Plane-dominated scenes: most inliers lie on one scene plane, so accepted
# seven-point samples are H-degenerate and DEGENSAC (fundamental_estimator.h:
# 341-572) replaces models through its nested plane-and-parallax GC-RANSAC.
# The oracle's trace counts its degenerate samples (ORACLE_TRACE_DEGENSAC).
@pytest.mark.parametrize("seed,plane", [(0, 0.8), (1, 0.8), (2, 0.9), (0, 0.9)])
def test_find_fundamental_degensac_matches_oracle(oracle, seed, plane):
    import pygcransac
    sc = synthetic.fundamental_scene(synthetic.scene_seed(7, seed), n=1500, inlier_ratio=0.5, plane_fraction=plane)
    p = oracle.params(1.0, 0.99, 0.975, 100000, 50, 0, 50, 8.0, seed)
    rc, Fo, mo, so = oracle.find_fundamental(sc["corr"], 480, 640, 480, 640, p)
    F, mask = pygcransac.findFundamentalMatrix(sc["corr"], 480, 640, 480, 640, None, threshold=1.0, conf=0.99,
                                               spatial_coherence_weight=0.975, max_iters=100000, min_iters=50,
                                               use_sprt=False, sampler=0, neighborhood=0, neighborhood_size=8,
                                               lo_number=50, seed=seed)
    st = pygcransac.last_stats()
    assert rc == mask.sum()
    assert np.array_equal(mask, mo)
    for key in ("iteration_number", "local_optimization_number", "graph_cut_number"):
        assert st[key] == so[key], key
    np.testing.assert_allclose(F, Fo, rtol=1e-12, atol=1e-14)


def test_degensac_repeat_calls_identical(oracle):
    # the nested run's labelling (spatial weight 0) on a pair whose grid has
    # edges must not read edge capacities left by earlier calls
    import pygcransac
    sc = synthetic.fundamental_scene(synthetic.scene_seed(7, 0), n=1500, inlier_ratio=0.5, plane_fraction=0.8)
    p = oracle.params(1.0, 0.99, 0.975, 100000, 50, 0, 50, 8.0, 0)
    rc, Fo, mo, so = oracle.find_fundamental(sc["corr"], 480, 640, 480, 640, p)
    for _ in range(3):
        F, mask = pygcransac.findFundamentalMatrix(sc["corr"], 480, 640, 480, 640, None, threshold=1.0, conf=0.99,
                                                   spatial_coherence_weight=0.975, max_iters=100000, min_iters=50,
                                                   use_sprt=False, sampler=0, neighborhood=0, neighborhood_size=8,
                                                   lo_number=50, seed=0)
        assert np.array_equal(mask, mo)
        assert pygcransac.last_stats()["iteration_number"] == so["iteration_number"]
        np.testing.assert_allclose(F, Fo, rtol=1e-12, atol=1e-14)


def test_nonminimal_fit_beyond_lds_bit_exact(oracle, engine):
    # 7,200 inliers: the fit's rows exceed one workgroup's LDS, so the first
    # rows stay in LDS and the rest in the fit's scratch (four-wave fit)
    sc = synthetic.fundamental_scene(synthetic.scene_seed(5, 1), n=12000, inlier_ratio=0.6)
    pts = sc["corr"]
    inl = np.nonzero(sc["gt"])[0]
    assert len(inl) > 4000
    w = np.zeros(len(pts))
    w[inl] = np.random.default_rng(3).uniform(0.0, 1.0, len(inl)) ** 2
    F = sc["F"] / np.linalg.norm(sc["F"])
    for sample, wt, wm, norm in [(inl, w, 1, 1), (inl, F.ravel(), 2, 0), (inl, None, 0, 1)]:
        om = oracle.fit_nonminimal(1, pts, sample, wt, wm, norm, 1.0)
        dm = engine.fit_nonminimal(1, pts, sample.astype(np.int32), wt, wm, norm, 1.0)
        assert om.shape == dm.shape
        assert np.array_equal(_bits(om), _bits(dm)), (wm, np.abs(om - dm).max())


def test_find_fundamental_large_scene_matches_oracle(oracle):
    import pygcransac
    sc = synthetic.fundamental_scene(synthetic.scene_seed(5, 2), n=10000, inlier_ratio=0.5)
    p = oracle.params(1.0, 0.99, 0.975, 100000, 50, 0, 50, 8.0, 3)
    rc, Fo, mo, so = oracle.find_fundamental(sc["corr"], 480, 640, 480, 640, p)
    Fd, md = pygcransac.findFundamentalMatrix(sc["corr"], 480, 640, 480, 640, None, threshold=1.0, conf=0.99,
                                              spatial_coherence_weight=0.975, max_iters=100000, min_iters=50,
                                              use_sprt=False, sampler=0, neighborhood=0, neighborhood_size=8,
                                              lo_number=50, seed=3)
    sd = pygcransac.last_stats()
    assert rc == md.sum() and rc > 4000
    assert np.array_equal(mo, md)
    for key in ("iteration_number", "local_optimization_number", "graph_cut_number"):
        assert so[key] == sd[key], key
    np.testing.assert_allclose(Fd, Fo, rtol=1e-12, atol=1e-14)


def test_find_fundamental_no_structure_matches_oracle(oracle):
    # random correspondences: the oracle's outcome (usually no model) and the
    # engine's agree, including the identity model and empty mask of the
    # driver when nothing is found
    import pygcransac
    rng = np.random.default_rng(9)
    corr = np.column_stack([rng.uniform(0, 640, 400), rng.uniform(0, 480, 400),
                            rng.uniform(0, 640, 400), rng.uniform(0, 480, 400)])
    p = oracle.params(0.05, 0.99, 0.975, 2000, 50, 0, 50, 8.0, 4)
    rc, Fo, mo, so = oracle.find_fundamental(corr, 480, 640, 480, 640, p)
    Fd, md = pygcransac.findFundamentalMatrix(corr, 480, 640, 480, 640, None, threshold=0.05, conf=0.99,
                                              spatial_coherence_weight=0.975, max_iters=2000, min_iters=50,
                                              use_sprt=False, sampler=0, neighborhood=0, neighborhood_size=8,
                                              lo_number=50, seed=4)
    assert rc == md.sum()
    assert np.array_equal(mo, md)
    assert pygcransac.last_stats()["iteration_number"] == so["iteration_number"]
    if rc == 0:
        assert Fd is None
    else:
        np.testing.assert_allclose(Fd, Fo, rtol=1e-12, atol=1e-14)
