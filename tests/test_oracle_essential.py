"""Pinning the essential-matrix oracle (CPU).

* Sturm roots: bit-identical to the reference's own root finder on the
  committed fixtures (tests/golden/sturm_fixtures.npz, written by
  tests/golden/make_sturm_fixtures.py from inc/maths/sturm_polynomial_solver.h),
  and equal to numpy's real roots on well-separated polynomials.
* Five-point minimal solver: exact recovery of a known E from five noise-free
  correspondences; every returned model satisfies det E = 0,
  2 E E^T E - tr(E E^T) E = 0 and the epipolar constraint on its sample.
* Non-minimal solver: exact recovery from 5, 8 and 60 noise-free matches,
  weighted and unweighted.
* Pose decomposition and bundle adjustment: the cheirality test keeps the true
  pose; the refined models contain the true E on exact data; on noisy data the
  weighted truncated Sampson cost does not increase.
"""
import os

import numpy as np
import pytest

from pygcransac import synthetic

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rot(axis, ang):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K


def _exact_pair(seed, n):
    rng = np.random.default_rng(seed)
    R = _rot(rng.normal(size=3), rng.uniform(0.05, 0.35))
    t = rng.normal(size=3)
    t /= np.linalg.norm(t)
    X = np.c_[rng.uniform(-1, 1, (n, 2)), rng.uniform(2, 8, n)]
    X2 = X @ R.T + t
    x1 = X[:, :2] / X[:, 2:]
    x2 = X2[:, :2] / X2[:, 2:]
    tx = np.array([[0, -t[2], t[1]], [t[2], 0, -t[0]], [-t[1], t[0], 0]])
    E = tx @ R
    return np.ascontiguousarray(np.c_[x1, x2]), E / np.linalg.norm(E), R, t


def _close_up_to_sign(A, B, tol):
    A = A.reshape(3, 3) / np.linalg.norm(A)
    B = B.reshape(3, 3) / np.linalg.norm(B)
    return min(np.abs(A - B).max(), np.abs(A + B).max()) < tol


def _invariants(E, tol):
    E = E.reshape(3, 3) / np.linalg.norm(E)
    assert abs(np.linalg.det(E)) < tol
    C = 2 * E @ E.T @ E - np.trace(E @ E.T) * E
    assert np.abs(C).max() < tol
    s = np.linalg.svd(E, compute_uv=False)
    assert abs(s[0] - s[1]) < 10 * tol and s[2] < 10 * tol


def test_sturm_matches_reference_fixtures(oracle):
    f = np.load(os.path.join(GOLDEN, "sturm_fixtures.npz"))
    for c, r, k in zip(f["coeffs"], f["roots"], f["counts"]):
        mine = oracle.sturm10(c)
        assert mine.size == k
        assert np.array_equal(mine.view(np.uint64), r[:k].view(np.uint64))


def test_sturm_against_numpy(oracle):
    rng = np.random.default_rng(3)
    for _ in range(20):
        r = np.sort(rng.uniform(-4, 4, 10))
        if np.min(np.diff(r)) < 0.05:
            continue
        mine = np.sort(oracle.sturm10(np.poly(r)[::-1].copy()))
        assert mine.size == 10
        np.testing.assert_allclose(mine, r, atol=1e-8)


@pytest.mark.parametrize("seed", range(6))
def test_five_point_exact_recovery(oracle, seed):
    pts, E, _, _ = _exact_pair(seed, 40)
    sample = np.arange(5, dtype=np.int32)[None, :]
    models, nm = oracle.e5_solve(pts, sample, orient=False)
    assert 1 <= nm[0] <= 10
    found = False
    for m in models[0, :nm[0]]:
        _invariants(m, 1e-6)
        for i in range(5):
            x1 = np.r_[pts[i, :2], 1.0]
            x2 = np.r_[pts[i, 2:], 1.0]
            assert abs(x2 @ m.reshape(3, 3) @ x1) < 1e-9
        found |= _close_up_to_sign(m, E, 1e-7)
    assert found


@pytest.mark.parametrize("n", [5, 8, 60])
def test_nonminimal_exact_recovery(oracle, n):
    pts, E, _, _ = _exact_pair(100 + n, max(n, 10))
    sample = np.arange(n)
    w = np.random.default_rng(n).uniform(0.2, 1.0, len(pts))
    for weights in (None, w):
        models = oracle.e_nonminimal(pts, sample, weights)
        assert len(models) >= 1
        for m in models:
            assert abs(np.linalg.norm(m) - 1.0) < 1e-12 and m[8] >= 0
        assert any(_close_up_to_sign(m, E, 1e-7) for m in models)


def test_motion_from_essential_keeps_true_pose(oracle):
    pts, E, R, t = _exact_pair(7, 20)
    x1 = np.r_[pts[0, :2], 1.0]
    x2 = np.r_[pts[0, 2:], 1.0]
    x1 /= np.linalg.norm(x1)
    x2 /= np.linalg.norm(x2)
    poses = oracle.motion_from_essential(E, x1, x2)
    assert len(poses) >= 1
    ok = False
    for p in poses:
        w, x, y, z = p[:4]
        Rq = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                       [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                       [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        ok |= np.abs(Rq - R).max() < 1e-9 and np.abs(p[4:] - t).max() < 1e-9
    assert ok


def _cost(pts, sample, w, E):
    E = E.reshape(3, 3)
    c = 0.0
    for k, i in enumerate(sample):
        x1 = np.r_[pts[i, :2], 1.0]
        x2 = np.r_[pts[i, 2:], 1.0]
        C = x2 @ E @ x1
        a = E[:2] @ x1
        b = E[:, :2].T @ x2
        c += w[k] * min(C * C / (a @ a + b @ b), 1.0)
    return c


def test_bundle_adjustment_exact_and_descending(oracle):
    pts, E, _, _ = _exact_pair(11, 80)
    sample = np.arange(80)
    init, out = oracle.e_bundle(pts, sample, np.ones(80))
    assert any(_close_up_to_sign(m, E, 1e-9) for m in out)
    noisy = pts + np.random.default_rng(5).normal(0, 1e-3, pts.shape)
    w = np.random.default_rng(6).uniform(0.5, 1.0, 80)
    init, out = oracle.e_bundle(noisy, sample, w)
    assert len(init) >= 1 and len(out) >= 1
    best_init = min(_cost(noisy, sample, w, m) for m in init)
    best_out = min(_cost(noisy, sample, w, m) for m in out)
    assert best_out <= best_init * (1 + 1e-12)


def test_find_essential_recovers_scene(oracle):
    sc = synthetic.essential_scene(synthetic.scene_seed(2, 0), n=2000, inlier_ratio=0.5)
    p = oracle.params(0.75, 0.99, 0.975, 100000, 50, 0, 50, 8.0, 3)
    rc, E, mask, st = oracle.find_essential(sc["corr"], sc["K1"], sc["K2"], 768, 1024, 768, 1024, p)
    rc2, E2, mask2, st2 = oracle.find_essential(sc["corr"], sc["K1"], sc["K2"], 768, 1024, 768, 1024, p)
    assert rc == rc2 and np.array_equal(mask, mask2) and np.array_equal(E, E2)
    gt = sc["gt"]
    assert (mask & gt).sum() >= 0.8 * gt.sum() and (mask & ~gt).sum() <= 0.02 * gt.sum()
    assert _close_up_to_sign(E, sc["E"], 5e-3)
