"""The oracle's radius neighbourhood (inc/neighborhood/flann_neighborhood_graph.h:
100-139) and the graph-cut labelling on it (inc/GCRANSAC.h:913-1021).  CPU
only.

What pins it:
- the neighbour lists against an exhaustive float32 search written here
  (numpy float32 arithmetic, no fused operations), including the order
  (distance, then index) and the reference's "drop the first entry" rule,
  which keeps a point in its own list when an equal point of lower index
  exists -- a restatement that dropped the point itself, compared in double,
  used < instead of <=, or compared against r instead of r*r fails here;
- points placed exactly on the radius, and a pair whose distance is above
  the radius in double but on it after rounding to float;
- the labelling against brute force over all 2^n labellings of the energy
  built here from the neighbour pairs: the SINK set must be the smallest
  minimiser's (the sink side Boykov-Kolmogorov leaves), and the graph's edges
  must come out in (p ascending, then p's list order) with p < q;
- a full run: no edge within the radius means the drivers return no model.
"""
import itertools

import numpy as np
import pytest

from pygcransac import synthetic


def _sqd32(P):
    f = P.astype(np.float32)
    d = f[:, None, :] - f[None, :, :]
    s = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]
    s = s + d[..., 2] * d[..., 2]
    s = s + d[..., 3] * d[..., 3]
    return s


def _lists(P, r):
    r2 = np.float32(r) * np.float32(r)
    D = _sqd32(P)
    out = []
    for i in range(len(P)):
        js = [j for j in range(len(P)) if D[i, j] <= r2]
        js.sort(key=lambda j: (D[i, j], j))
        out.append(js[1:])
    return out


def _flat(lists):
    off = np.zeros(len(lists) + 1, np.int64)
    off[1:] = np.cumsum([len(l) for l in lists])
    return off, np.array([j for l in lists for j in l], np.int32)


def test_lists_match_exhaustive_float_search(oracle):
    rng = np.random.default_rng(3)
    for trial in range(6):
        n = 300
        P = rng.uniform(0, 200, (n, 4))
        P[:20] = P[20:40] + rng.normal(0, 3, (20, 4))  # some close pairs
        P[50:53] = P[60]                             # exact duplicates
        r = [15.0, 25.0, 7.5, 40.0, 0.0, 12.25][trial]
        off, nbr = oracle.radius_neighbors(P, r)
        eoff, enbr = _flat(_lists(P, r))
        assert np.array_equal(off, eoff), trial
        assert np.array_equal(nbr, enbr), trial


def test_radius_boundary_and_float_rounding(oracle):
    P = np.array([[0, 0, 0, 0], [3, 4, 0, 0], [3, 4, 0, 1e-3], [0, 0, 1 + 1e-9, 0]], np.float64)
    off, nbr = oracle.radius_neighbors(P, 5.0)
    l0 = list(nbr[off[0]:off[1]])
    # (3,4,0,0) is exactly on the radius (25 <= 25); (3,4,0,1e-3) is beyond
    assert 1 in l0 and 2 not in l0
    # d = 1 + 1e-9 > 1 in double, but the float coordinate is 1.0: included
    off, nbr = oracle.radius_neighbors(P[[0, 3]], 1.0)
    assert list(nbr) == [1, 0]
    assert oracle.radius_sq_distance(P[0], P[3]) == 1.0


def test_equal_points_keep_later_copies_in_their_lists(oracle):
    P = np.array([[10, 10, 10, 10]] * 3 + [[11, 10, 10, 10]], np.float64)
    off, nbr = oracle.radius_neighbors(P, 2.0)
    lists = [list(nbr[off[i]:off[i + 1]]) for i in range(4)]
    assert lists[0] == [1, 2, 3]
    # the first entry of point 1's list is point 0 (distance 0, lower index):
    # it is dropped and point 1 stays in its own list
    assert lists[1] == [1, 2, 3]
    assert lists[2] == [1, 2, 3]
    assert lists[3] == [0, 1, 2]  # itself first (distance 0), then the copies at distance 1


def test_lists_are_symmetric_and_sorted(oracle):
    rng = np.random.default_rng(9)
    P = rng.uniform(0, 100, (400, 4))
    off, nbr = oracle.radius_neighbors(P, 18.0)
    D = _sqd32(P)
    S = set()
    for i in range(len(P)):
        l = nbr[off[i]:off[i + 1]]
        key = [(D[i, j], j) for j in l]
        assert key == sorted(key)
        for j in l:
            if j != i:
                S.add((min(i, j), max(i, j)))
    for (a, b) in S:
        assert b in nbr[off[a]:off[a + 1]] and a in nbr[off[b]:off[b + 1]]


def _energy_terms(P, model_r2, thr, lam, r):
    """Unary (E(0), E(1)) and pairwise (A, B, C, D) terms built from the
    GCRANSAC::labeling definition, on the exhaustive float radius pairs."""
    sq = thr * thr * 9 / 4
    d = np.clip(model_r2 / sq, 0, 1)
    near = model_r2 <= sq
    ua = np.where(near, (1 - lam) * (1 - d), 0.0)
    ub = np.where(near, 0.0, (1 - lam) * d)
    pairs = []
    for p, l in enumerate(_lists(P, r)):
        for q in l:
            if q > p:
                pairs.append((p, q, 0.5 * (d[p] + d[q]) * lam, lam, lam, 0.0))
    return ua, ub, pairs


def _brute(ua, ub, pairs):
    """Energies of all 2^n labellings (x = 1: SINK)."""
    n = len(ua)
    X = np.array(list(itertools.product((0, 1), repeat=n)), np.int8)
    E = np.where(X == 1, ub[None, :], ua[None, :]).sum(1)
    for (p, q, A, B, C, D) in pairs:
        tab = np.array([A, B, C, D])
        E = E + tab[2 * X[:, p] + X[:, q]]
    return X, E


def test_labelling_on_radius_graph_is_the_smallest_minimiser(oracle):
    """The SINK set minimises the energy to 1e-9 relative.  Where no other
    labelling comes within that of the minimum it is the minimiser itself;
    where several do (a far point whose k SINK neighbours cost k * lambda =
    1 - lambda, say), they differ in the last bits of the FP64 capacities,
    which Boykov-Kolmogorov's FP64 arithmetic resolves either way, so only
    the energy is checked."""
    rng = np.random.default_rng(11)
    checked = unique = 0
    for trial in range(30):
        n = 12
        sc = synthetic.homography_scene(synthetic.scene_seed(12, trial), n=200, inlier_ratio=0.5)
        # a cluster of points so the radius graph has edges
        P = sc["corr"][:n].copy()
        P[:, :2] = P[0, :2] + rng.uniform(-15, 15, (n, 2))
        P[:, 2:] = P[0, 2:] + rng.uniform(-15, 15, (n, 2))
        H = sc["H"]
        x = np.c_[P[:, :2], np.ones(n)] @ H.T
        # move points off the model by graded amounts
        P[:, 2:] = x[:, :2] / x[:, 2:] + rng.normal(0, 1, (n, 2)) * rng.uniform(0, 6, (n, 1))
        thr, lam, r = 3.0, float(rng.choice([0.975, 0.5, 0.2])), float(rng.uniform(10, 30))
        mine, g = oracle.labeling_radius(0, P, H, lam, thr, r, want_graph=True)
        xh = np.c_[P[:, :2], np.ones(n)] @ H.T
        r2 = ((xh[:, :2] / xh[:, 2:] - P[:, 2:]) ** 2).sum(1)
        ua, ub, pairs = _energy_terms(P, r2, thr, lam, r)
        # edges in (p ascending, list order), p < q, with the reference's
        # capacities B - A (p -> q) and C (q -> p)
        assert [(int(a), int(b)) for a, b in zip(g["ep"], g["eq"])] == [(p, q) for (p, q, *_ ) in pairs]
        np.testing.assert_allclose(g["cap_pq"], [B - A for (_, _, A, B, _, _) in pairs], rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(g["cap_qp"], [C for (*_, C, _) in pairs], rtol=0, atol=0)
        X, E = _brute(ua, ub, pairs)
        best = E.min()
        tol = 1e-9 * (1 + abs(best))
        e_mine = E[int("".join(str(int(v)) for v in mine), 2)]
        assert e_mine <= best + tol, trial
        near = X[E <= best + tol]
        if len(near) == 1:
            assert np.array_equal(mine, near[0].astype(bool)), trial
            unique += 1
        checked += len(pairs) > 0
    assert checked >= 20 and unique >= 15


def test_no_edge_within_radius_returns_no_model(oracle):
    sc = synthetic.homography_scene(synthetic.scene_seed(12, 99), n=300, inlier_ratio=0.6)
    # radius 1e-3: no two correspondences are that close
    p = oracle.params(3.0, 0.99, 0.975, 1000, 50, 0, 50, 1e-3, 1, neighborhood=1)
    rc, H, mask, st = oracle.find_homography(sc["corr"], 480, 640, 480, 640, p)
    assert rc == 0 and not mask.any() and st["iteration_number"] == 0
    # with a usable radius the same call finds the plane
    p = oracle.params(3.0, 0.99, 0.975, 1000, 50, 0, 50, 40.0, 1, neighborhood=1)
    rc, H, mask, st = oracle.find_homography(sc["corr"], 480, 640, 480, 640, p)
    assert rc > 0 and (mask & sc["gt"]).sum() >= 0.95 * sc["gt"].sum()


def test_essential_with_reference_defaults_recovers_pose(oracle):
    """findEssentialMatrix's defaults: spatial weight 0.1, radius graph of
    radius 20 px on the pixel coordinates."""
    sc = synthetic.essential_scene(synthetic.scene_seed(12, 7), n=1500, inlier_ratio=0.4)
    p = oracle.params(0.75, 0.99, 0.1, 10000, 50, 0, 50, 20.0, 3, neighborhood=1)
    rc, E, mask, st = oracle.find_essential(sc["corr"], sc["K1"], sc["K2"], 768, 1024, 768, 1024, p)
    assert rc > 0
    # 0.75 px noise at a 0.75 px threshold: about 85 % of the generator's
    # inliers lie inside the threshold of the true pose
    assert (mask & sc["gt"]).sum() >= 0.8 * sc["gt"].sum()
    assert st["graph_cut_number"] > 0
