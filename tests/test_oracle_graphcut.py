"""The oracle's graph-cut labelling against the reference's own max-flow.

Golden vectors come from the reference's Boykov-Kolmogorov code (energy.h /
graph.h / maxflow.cpp, compiled from the reference sources by
tests/golden/make_bk_fixtures.py).  Brute force over all 2^n labellings pins
the minimum energy independently of both max-flow codes.
"""
import itertools
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "bk_fixtures.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def _energy(t, x):
    e = 0.0
    for i in range(t["ua"].size):
        e += t["ub"][i] if x[i] else t["ua"][i]
    for k in range(t["ep"].size):
        a, b = x[t["ep"][k]], x[t["eq"][k]]
        e += (t["ta"][k], t["tb"][k], t["tc"][k], t["td"][k])[2 * a + b]
    return e


def test_energy_usage_example(gold):
    # energy.h:31-69: E(x,y,z) = x - 2y + 3(1-z) - 4xy + 5|y-z|, minimum -5 at (1, 1, 1)
    assert gold["example_energy"][0] == -5.0
    assert list(gold["example_xyz"]) == [1, 1, 1]


def test_random_energies_match_reference_and_brute_force(oracle, gold):
    for r in range(40):
        t = {k: gold["rand_%d_%s" % (r, k)] for k in ("ua", "ub", "ep", "eq", "ta", "tb", "tc", "td")}
        ref = gold["rand_%d_sink" % r]
        mine = oracle.energy_mincut(t)
        assert np.array_equal(mine, ref), r
        n = t["ua"].size
        best = min(_energy(t, x) for x in itertools.product((0, 1), repeat=n))
        assert abs(_energy(t, mine.astype(int)) - best) <= 1e-9 * (1 + abs(best)), r
        assert abs(gold["rand_%d_energy" % r][0] - best) <= 1e-9 * (1 + abs(best)), r


def test_scene_labelings_match_reference(oracle, gold):
    eps = np.finfo(np.float64).eps
    cell = np.array([(640 + eps) / 8, (480 + eps) / 8, (640 + eps) / 8, (480 + eps) / 8])
    for name in gold["scene_names"]:
        seed = int(str(name).split("_")[1])
        pts = gold["scene_%d_pts" % seed]
        model = gold[str(name) + "_model"]
        ref = gold[str(name) + "_sink"]
        mine = oracle.labeling(0, pts, model, 0.975, 3.0, cell, 8)
        assert np.array_equal(mine, ref), name
        t = oracle.energy_terms(0, pts, model, 0.975, 3.0, cell, 8)
        assert np.array_equal(oracle.energy_mincut(t), ref), name


def test_sink_set_is_minimal_cut_side(oracle):
    # A chain s -> 0 -> 1 -> t with equal capacities has two minimum cuts; the
    # reference labels the nodes that still reach the sink, i.e. none here.
    t = dict(ua=np.array([0.0, 1.0]), ub=np.array([1.0, 0.0]), ep=np.array([0], np.int32), eq=np.array([1], np.int32),
             ta=np.array([0.0]), tb=np.array([1.0]), tc=np.array([0.0]), td=np.array([0.0]))
    sink = oracle.energy_mincut(t)
    assert _energy(t, sink.astype(int)) == 1.0
