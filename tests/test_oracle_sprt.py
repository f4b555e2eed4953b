"""SPRT preemptive verification in the oracle (inc/preemption/preemption_sprt.h:
206-379).  CPU only.

What pins it:
- A, the decision threshold: the reference runs at most ten steps of the
  fixed-point iteration A <- K + ln A from A = K, K = t_M C / m_S + 1,
  C = (1-d) ln((1-d)/(1-e)) + d ln(d/e) (Wald's test, t_M = 1/m); K is
  close to 1, where the iteration converges slowly, so A lies between K and
  the root of A = K + ln A (scipy's root finder), and equals the tenth
  iterate computed here -- a swapped d / e, a sign error or a dropped term
  fails;
- the random pool: a permutation built with rand() % max and a swap to the
  end, replayed here with the C library's own rand();
- the test: points in pool order from the running pool index, rejection at
  the first prefix whose likelihood ratio (d/e)^I ((1-d)/(1-e))^(T-I) exceeds
  A (evaluated here in log space, from residuals computed here), and the next
  model starting where the last one stopped;
- the history rules: after an accepted model scoring above the best,
  e = inliers / N; after a rejection, d = tested inliers / N when that moves
  d by more than 5 %;
- the score: the tested inliers' count and sum of 1 - r^2 / tau^2.
And at run level: GC-RANSAC with SPRT on still recovers the model.
"""
import ctypes
import ctypes.util
import math

import numpy as np
import pytest
from scipy.optimize import brentq

from pygcransac import synthetic


def _C(e, d):
    return (1 - d) * math.log((1 - d) / (1 - e)) + d * math.log(d / e)


@pytest.mark.parametrize("e,d,m,ms", [(0.01, 0.001, 4, 1), (1e-5, 1e-6, 7, 3), (0.3, 0.02, 5, 10), (0.5, 0.1, 7, 3),
                                      (0.05, 0.00025, 4, 1)])
def test_threshold_follows_the_wald_iteration(oracle, e, d, m, ms):
    K = (1.0 / m) * _C(e, d) / ms + 1
    root = brentq(lambda A: A - K - math.log(A), 1.0 + 1e-12, K + 50)
    it, prev = K, K
    for _ in range(10):
        it = K + math.log(prev)
        if abs(it - prev) < 1.5e-8:
            break
        prev = it
    A = oracle.sprt_threshold_A(e, d, m, ms)
    assert K < A <= root * (1 + 1e-12)
    assert A == it


def _homography_models(sc, rng, k):
    H = sc["H"]
    models = []
    for j in range(k):
        r = rng.uniform()
        if r < 0.2:
            M = H.copy()
        elif r < 0.5:
            M = H + rng.normal(0, 1e-4, (3, 3)) * np.abs(H)
        else:
            M = rng.normal(size=(3, 3))
        models.append(M / M[2, 2])
    return np.array(models)


def _h_r2(pts, M):
    x = np.c_[pts[:, :2], np.ones(len(pts))] @ M.T
    return ((x[:, :2] / x[:, 2:] - pts[:, 2:]) ** 2).sum(1)


def test_pool_is_the_c_library_shuffle(oracle):
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    sc = synthetic.homography_scene(synthetic.scene_seed(14, 0), n=500, inlier_ratio=0.5)
    out = oracle.sprt_sequence(0, sc["corr"], sc["H"][None], 4.5, 0.01, 4, 1, 77, [0.0])
    libc.srand(ctypes.c_uint(77))
    pool = list(range(500))
    mx = 500
    for _ in range(500):
        r = libc.rand() % mx
        mx -= 1
        pool[r], pool[mx] = pool[mx], pool[r]
    assert list(out["pool"]) == pool
    assert sorted(pool) == list(range(500))


def test_test_decisions_follow_walds_likelihood_ratio(oracle):
    rng = np.random.default_rng(5)
    checked = rejected = accepted = 0
    for trial in range(6):
        sc = synthetic.homography_scene(synthetic.scene_seed(14, 10 + trial), n=800, inlier_ratio=0.4)
        pts = sc["corr"]
        n = len(pts)
        models = _homography_models(sc, rng, 40)
        tau = 1.5 * 3.0
        ratio = [0.01, 0.05, 0.2][trial % 3]
        # a best value that rises as the sequence goes on
        best = np.linspace(0, 0.3 * n, len(models))
        out = oracle.sprt_sequence(0, pts, models, tau, ratio, 4, 1, 1000 + trial, best)
        pool = out["pool"]
        e, d = ratio, ratio / 10.0
        A = oracle.sprt_threshold_A(e, d, 4, 1)
        start = 0
        for j, M in enumerate(models):
            r2 = _h_r2(pts, M)
            if np.any(np.abs(r2 - tau * tau) < 1e-6 * tau * tau):
                break  # a residual on the threshold: stop comparing this sequence
            assert out["start"][j] == start, (trial, j)
            order = pool[(start + np.arange(n)) % n]
            inl = r2[order] < tau * tau
            # log-likelihood ratio after each tested point
            llr = np.cumsum(np.where(inl, math.log(d / e), math.log((1 - d) / (1 - e))))
            over = np.nonzero(llr > math.log(A))[0]
            if len(over) and (llr[over[0]] - math.log(A)) < 1e-9:
                break  # a ratio on the threshold: stop comparing this sequence
            T = over[0] + 1 if len(over) else n
            assert out["accepted"][j] == (len(over) == 0), (trial, j)
            assert out["tested"][j] == T, (trial, j)
            I = int(inl[:T].sum())
            assert out["count"][j] == I
            expect = np.sum(1.0 - r2[order[:T]][inl[:T]] / (tau * tau))
            assert abs(out["value"][j] - expect) <= 1e-9 * max(1.0, abs(expect))
            # history rules
            if len(over) == 0:
                accepted += 1
                if out["value"][j] > best[j]:
                    e = I / n
                    A = oracle.sprt_threshold_A(e, d, 4, 1)
            else:
                rejected += 1
                de = I / n
                if de > 0 and abs(d - de) / d > 0.05:
                    d = de
                    A = oracle.sprt_threshold_A(e, d, 4, 1)
            assert out["eps"][j] == pytest.approx(e, rel=1e-15) and out["delta"][j] == pytest.approx(d, rel=1e-15)
            assert out["A"][j] == pytest.approx(A, rel=1e-15)
            start = (start + T) % n
            checked += 1
    assert checked >= 150 and rejected >= 30 and accepted >= 10


@pytest.mark.parametrize("kind", ["homography", "fundamental", "essential"])
def test_runs_with_sprt_recover_the_model(oracle, kind):
    rng_seed = {"homography": 1, "fundamental": 2, "essential": 3}[kind]
    if kind == "homography":
        sc = synthetic.homography_scene(synthetic.scene_seed(14, 50), n=2000, inlier_ratio=0.3)
        p = oracle.params(3.0, 0.99, 0.975, 10000, 50, 0, 50, 8.0, rng_seed, use_sprt=True)
        rc, M, mask, st = oracle.find_homography(sc["corr"], 480, 640, 480, 640, p)
    elif kind == "fundamental":
        sc = synthetic.fundamental_scene(synthetic.scene_seed(14, 51), n=2000, inlier_ratio=0.4)
        p = oracle.params(1.0, 0.99, 0.0, 10000, 50, 0, 50, 8.0, rng_seed, use_sprt=True,
                          min_inlier_ratio_for_sprt=1e-5)
        rc, M, mask, st = oracle.find_fundamental(sc["corr"], 480, 640, 480, 640, p)
    else:
        sc = synthetic.essential_scene(synthetic.scene_seed(14, 52), n=2000, inlier_ratio=0.4)
        p = oracle.params(0.75, 0.99, 0.1, 10000, 50, 0, 50, 20.0, rng_seed, neighborhood=1, use_sprt=True,
                          min_inlier_ratio_for_sprt=0.01)
        rc, M, mask, st = oracle.find_essential(sc["corr"], sc["K1"], sc["K2"], 768, 1024, 768, 1024, p)
    assert rc > 0
    assert (mask & sc["gt"]).sum() >= 0.8 * sc["gt"].sum()
