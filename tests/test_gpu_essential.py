"""GPU parity of the essential-matrix path against the CPU oracle.

Bar: bit-exact minimal-solver statuses and models, non-minimal fits and the
driver's pose bundle adjustment with the oracle's bits (identical IEEE
operation sequences, no fused multiply-add, identical reduction orders), and
full findEssentialMatrix runs with identical inlier masks and statistics
(model within 1e-12).
"""
import numpy as np
import pytest

from pygcransac import synthetic

pytestmark = pytest.mark.gpu

THR = 0.75
F_PX = 1000.0


def _scene(seed, n=2000, eps=0.5):
    return synthetic.essential_scene(synthetic.scene_seed(2, seed), n=n, inlier_ratio=eps)


def _bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def _normalised(oracle, sc):
    return oracle.normalise_correspondences(sc["corr"], sc["K1"], sc["K2"])


@pytest.mark.parametrize("seed", [0, 1])
def test_five_point_bit_exact(oracle, engine, seed):
    sc = _scene(seed)
    pts = _normalised(oracle, sc)
    rng = np.random.default_rng(seed)
    samples = np.stack([rng.choice(len(pts), 5, replace=False) for _ in range(1500)]).astype(np.int32)
    om, onm = oracle.e5_solve(pts, samples, orient=True)
    dm, dst = engine.solve_minimal(2, pts, samples)
    dnm = (dst == 2).sum(axis=1)
    assert np.array_equal(onm, dnm)
    assert onm.sum() > 1000
    for k in range(len(samples)):
        n = onm[k]
        assert np.array_equal(_bits(om[k, :n]), _bits(dm[k, :n])), k


@pytest.mark.parametrize("mode", ["lo", "five", "irls"])
def test_nonminimal_fit_bit_exact(oracle, engine, mode):
    sc = _scene(3)
    pts = _normalised(oracle, sc)
    inl = np.nonzero(sc["gt"])[0]
    rng = np.random.default_rng(7)
    if mode == "lo":
        cases = [(rng.choice(inl, 35, replace=False), None) for _ in range(6)]
    elif mode == "five":
        cases = [(rng.choice(inl, 5, replace=False), None) for _ in range(6)]
    else:
        w = np.zeros(len(pts))
        w[inl] = rng.uniform(0.0, 1.0, len(inl)) ** 2
        cases = [(inl[:300], w), (inl, w)]
    for sample, w in cases:
        om = oracle.e_nonminimal(pts, sample, w)
        dm = engine.fit_nonminimal(2, pts, sample.astype(np.int32), w, 0 if w is None else 1, 0, THR / F_PX)
        assert om.shape == dm.shape
        assert np.array_equal(_bits(om), _bits(dm)), np.abs(om - dm).max()


def test_pose_bundle_adjustment_bit_exact(oracle, engine):
    sc = _scene(4)
    pts = _normalised(oracle, sc)
    inl = np.nonzero(sc["gt"])[0]
    E = sc["E"] / np.linalg.norm(sc["E"])
    t = THR / F_PX
    for sample in (inl[:200], inl):
        om = oracle.e_post_bundle(pts, sample, E, 1.5 * t)
        dm = engine.fit_nonminimal(2, pts, sample.astype(np.int32), E.ravel(), 3, 0, t)
        assert om.shape == dm.shape and len(om) >= 1
        assert np.array_equal(_bits(om), _bits(dm)), np.abs(om - dm).max()


@pytest.mark.parametrize("seed,n,eps", [(0, 2000, 0.5), (1, 3000, 0.35), (2, 1500, 0.6)])
def test_find_essential_matches_oracle(oracle, seed, n, eps):
    import pygcransac
    sc = _scene(seed, n, eps)
    corr = sc["corr"]
    p = oracle.params(THR, 0.99, 0.975, 100000, 50, 0, 50, 8.0, seed)
    rc, Eo, mo, so = oracle.find_essential(corr, sc["K1"], sc["K2"], 768, 1024, 768, 1024, p)
    Ed, md = pygcransac.findEssentialMatrix(corr, sc["K1"], sc["K2"], 768, 1024, 768, 1024, None, threshold=THR,
                                            conf=0.99, spatial_coherence_weight=0.975, max_iters=100000,
                                            min_iters=50, use_sprt=False, sampler=0, neighborhood=0,
                                            neighborhood_size=8, lo_number=50, seed=seed)
    sd = pygcransac.last_stats()
    assert rc == md.sum()
    assert np.array_equal(mo, md)
    for key in ("iteration_number", "local_optimization_number", "graph_cut_number"):
        assert so[key] == sd[key], key
    np.testing.assert_allclose(Ed, Eo, rtol=1e-12, atol=1e-14)
    assert (md & sc["gt"]).sum() >= 0.8 * sc["gt"].sum()


def test_batch_with_essential_pairs(oracle):
    import pygcransac
    probs = []
    for i, (kind, n, eps) in enumerate([("essential", 1200, 0.5), ("homography", 900, 0.5),
                                        ("essential", 2500, 0.4), ("fundamental", 1000, 0.6)]):
        if kind == "essential":
            sc = _scene(10 + i, n, eps)
            probs.append(dict(kind=kind, correspondences=sc["corr"], K1=sc["K1"], K2=sc["K2"], h1=768, w1=1024,
                              h2=768, w2=1024, threshold=THR))
        elif kind == "homography":
            sc = synthetic.homography_scene(synthetic.scene_seed(3, 20 + i), n=n, inlier_ratio=eps)
            probs.append(dict(kind=kind, correspondences=sc["corr"], h1=480, w1=640, h2=480, w2=640, threshold=3.0))
        else:
            sc = synthetic.fundamental_scene(synthetic.scene_seed(3, 20 + i), n=n, inlier_ratio=eps)
            probs.append(dict(kind=kind, correspondences=sc["corr"], h1=480, w1=640, h2=480, w2=640, threshold=1.0))
        probs[-1].update(conf=0.99, spatial_coherence_weight=0.975, max_iters=100000, min_iters=50,
                         use_sprt=False, sampler=0, neighborhood=0, neighborhood_size=8.0, lo_number=50, seed=50 + i)
    res = pygcransac.findBatch(probs)
    stats = pygcransac.last_batch_stats()
    for i, pb in enumerate(probs):
        p = oracle.params(pb["threshold"], 0.99, 0.975, 100000, 50, 0, 50, 8.0, pb["seed"])
        if pb["kind"] == "essential":
            rc, Mo, mo, so = oracle.find_essential(pb["correspondences"], pb["K1"], pb["K2"], 768, 1024, 768, 1024, p)
        elif pb["kind"] == "homography":
            rc, Mo, mo, so = oracle.find_homography(pb["correspondences"], 480, 640, 480, 640, p)
        else:
            rc, Mo, mo, so = oracle.find_fundamental(pb["correspondences"], 480, 640, 480, 640, p)
        M, mask = res[i]
        assert np.array_equal(mask, mo), i
        for key in ("iteration_number", "local_optimization_number", "graph_cut_number"):
            assert stats[i][key] == so[key], (i, key)
        np.testing.assert_allclose(M, Mo, rtol=1e-12, atol=1e-14, err_msg=str(i))
