"""Batched engine runs (many pairs of uneven size and mixed problem kinds in
one gcr_find_batch call) against the CPU oracle run pair by pair.

Pairs in one batch suspend at different points of their loops (one in a
local optimisation while another asks for its next candidate chunk), so
this checks that no pair reads another pair's device state.  Bar: identical
inlier masks and statistics per pair, models within 1e-12.
"""
import numpy as np
import pytest

from pygcransac import synthetic

pytestmark = pytest.mark.gpu

# (kind, n, inlier ratio)
PAIRS = [("homography", 400, 0.35), ("fundamental", 1500, 0.5), ("homography", 2600, 0.6),
         ("fundamental", 900, 0.6), ("homography", 1200, 0.45), ("fundamental", 2200, 0.45),
         ("homography", 2000, 0.5), ("fundamental", 1200, 0.55)]


def _problems():
    out = []
    for i, (kind, n, eps) in enumerate(PAIRS):
        if kind == "homography":
            sc = synthetic.homography_scene(synthetic.scene_seed(3, i), n=n, inlier_ratio=eps)
            thr = 3.0
        else:
            sc = synthetic.fundamental_scene(synthetic.scene_seed(3, i), n=n, inlier_ratio=eps)
            thr = 1.0
        out.append(dict(kind=kind, correspondences=sc["corr"], h1=480, w1=640, h2=480, w2=640, threshold=thr,
                        conf=0.99, spatial_coherence_weight=0.975, max_iters=100000, min_iters=50, use_sprt=False,
                        sampler=0, neighborhood=0, neighborhood_size=8.0, lo_number=50, seed=100 + i))
    return out


def test_batch_matches_oracle_pair_by_pair(oracle):
    import pygcransac
    probs = _problems()
    res = pygcransac.findBatch(probs)
    stats = pygcransac.last_batch_stats()
    for i, pb in enumerate(probs):
        p = oracle.params(pb["threshold"], 0.99, 0.975, 100000, 50, 0, 50, 8.0, pb["seed"])
        if pb["kind"] == "homography":
            rc, Mo, mo, so = oracle.find_homography(pb["correspondences"], 480, 640, 480, 640, p)
        else:
            rc, Mo, mo, so = oracle.find_fundamental(pb["correspondences"], 480, 640, 480, 640, p)
        M, mask = res[i]
        case = "pair %d (%s, n=%d)" % (i, pb["kind"], len(mask))
        assert np.array_equal(mask, mo), case
        for key in ("iteration_number", "local_optimization_number", "graph_cut_number"):
            assert stats[i][key] == so[key], (case, key)
        np.testing.assert_allclose(M, Mo, rtol=1e-12, atol=1e-14, err_msg=case)


def test_batch_repeatable(oracle):
    import pygcransac
    probs = _problems()[:4]
    a = pygcransac.findBatch(probs)
    b = pygcransac.findBatch(probs)
    for (Ma, ma), (Mb, mb) in zip(a, b):
        assert np.array_equal(ma, mb)
        assert np.array_equal(Ma.view(np.uint64), Mb.view(np.uint64))
