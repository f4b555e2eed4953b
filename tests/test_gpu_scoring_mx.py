"""Candidate scoring on the matrix cores (FP16 hi/lo MFMA rejection of proven
outliers, FP64 in point order for every other point) against the FP64
scoring kernel: whole GC-RANSAC runs must be bit-identical -- masks, models,
iteration / LO / graph-cut counts -- since every accepted model, every MSAC
value and every validity count comes from candidate scores.

The batches cover homography (transfer error), fundamental (Sampson, with
DEGENSAC's plane-and-parallax estimator on plane-dominated scenes) and
essential pairs (Sampson on K-normalised points), from 15 % to 60 % inliers.
"""
import numpy as np
import pytest

from pygcransac import synthetic

pytestmark = pytest.mark.gpu


def _problems():
    out = []
    spec = [("homography", 1800, 0.5), ("fundamental", 3000, 0.3), ("essential", 2500, 0.4),
            ("fundamental", 4000, 0.15), ("homography", 900, 0.2), ("fundamental", 2000, 0.6),
            ("essential", 4000, 0.25), ("fundamental_plane", 2500, 0.5), ("homography", 6000, 0.35),
            ("fundamental", 800, 0.45)]
    for i, (kind, n, eps) in enumerate(spec):
        seed = synthetic.scene_seed(7, 500 + i)
        base = dict(conf=0.99, spatial_coherence_weight=0.975, max_iters=20000, min_iters=50, use_sprt=False,
                    sampler=0, neighborhood=0, neighborhood_size=8.0, lo_number=50, seed=900 + i)
        if kind == "homography":
            sc = synthetic.homography_scene(seed, n=n, inlier_ratio=eps)
            base.update(kind=kind, correspondences=sc["corr"], h1=480, w1=640, h2=480, w2=640, threshold=3.0)
        elif kind == "essential":
            sc = synthetic.essential_scene(seed, n=n, inlier_ratio=eps)
            base.update(kind=kind, correspondences=sc["corr"], K1=sc["K1"], K2=sc["K2"], h1=768, w1=1024, h2=768,
                        w2=1024, threshold=0.75)
        else:
            sc = synthetic.fundamental_scene(seed, n=n, inlier_ratio=eps,
                                             plane_fraction=0.9 if kind == "fundamental_plane" else 0.0)
            base.update(kind="fundamental", correspondences=sc["corr"], h1=480, w1=640, h2=480, w2=640,
                        threshold=1.0)
        out.append(base)
    return out


def _run(monkeypatch, probs, mfma):
    import pygcransac
    monkeypatch.setenv("GCR_SCORE_MFMA", "1" if mfma else "0")
    res = pygcransac.findBatch(probs)
    return res, pygcransac.last_batch_stats()


def test_matrix_core_scoring_identical_to_fp64(monkeypatch):
    probs = _problems()
    print("scoring: FP64 kernel", flush=True)
    r64, s64 = _run(monkeypatch, probs, False)
    print("scoring: matrix-core kernel", flush=True)
    rmx, smx = _run(monkeypatch, probs, True)
    for i, ((Ma, ma), (Mb, mb)) in enumerate(zip(r64, rmx)):
        case = "pair %d (%s)" % (i, probs[i]["kind"])
        assert np.array_equal(ma, mb), case
        assert (Ma is None) == (Mb is None), case
        if Ma is not None:
            assert np.array_equal(Ma.view(np.uint64), Mb.view(np.uint64)), case
        for key in ("iteration_number", "local_optimization_number", "graph_cut_number", "num_inliers",
                    "models_scored"):
            assert s64[i][key] == smx[i][key], (case, key)
        assert s64[i]["score"] == smx[i]["score"], case
    assert sum(m.sum() for _, m in rmx) > 5000


def test_matrix_core_scoring_low_inlier_ratio(monkeypatch):
    """4,000 correspondences at 5 % inliers (most candidate models are far
    from every point): identical runs at the API default cap."""
    probs = []
    for i in range(6):
        sc = synthetic.fundamental_scene(synthetic.scene_seed(4, 3000 + i), n=4000, inlier_ratio=0.05)
        probs.append(dict(kind="fundamental", correspondences=sc["corr"], h1=480, w1=640, h2=480, w2=640,
                          threshold=1.0, conf=0.99, spatial_coherence_weight=0.975, max_iters=10000, min_iters=50,
                          use_sprt=False, sampler=0, neighborhood=0, neighborhood_size=8.0, lo_number=50,
                          seed=40 + i))
    r64, s64 = _run(monkeypatch, probs, False)
    rmx, smx = _run(monkeypatch, probs, True)
    for i, ((Ma, ma), (Mb, mb)) in enumerate(zip(r64, rmx)):
        assert np.array_equal(ma, mb), i
        for key in ("iteration_number", "local_optimization_number", "graph_cut_number", "num_inliers"):
            assert s64[i][key] == smx[i][key], (i, key)
