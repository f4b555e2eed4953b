"""Host-side checks of bench.py: the one-GPU contract and the workload
generators (no GPU needed)."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_mixed_workload_pairs():
    import bench
    sc = bench.make_pairs("mixed", 256)
    kinds = np.array([s["kind"] for s in sc])
    ns = np.array([len(s["corr"]) for s in sc])
    eps = np.array([s["gt"].mean() for s in sc])
    assert (kinds[0::2] == 0).all() and (kinds[1::2] == 1).all()
    assert ns.min() >= 200 and ns.max() < 8000
    # log-uniform counts: the median sits near sqrt(200 * 8000) = 1265
    assert 900 < np.median(ns) < 1800
    assert eps.min() > 0.08 and eps.max() < 0.72
    assert all(s["thr"] == (3.0 if s["kind"] == 0 else 1.0) for s in sc)
    again = bench.make_pairs("mixed", 4)
    assert all(np.array_equal(a["corr"], b["corr"]) for a, b in zip(again, sc[:4]))


def test_guided_workload_pairs_are_the_held_out_pairs():
    import bench
    from pygcransac import train_guided
    sc = bench.make_pairs("guided", 3)
    probs, gts = train_guided.held_out(3)
    for s, p, g in zip(sc, probs, gts):
        assert np.array_equal(s["corr"], p["correspondences"])
        assert np.array_equal(s["gt"], g)
        assert abs(g.mean() - 0.05) < 1e-9
