"""The oracle's DEGENSAC branch (fundamental_estimator.h:341-572) on
plane-dominated scenes: the degenerate-sample test fires (counted through
the oracle's trace), and the branch keeps what the mathematics fixes: the
returned model is a fundamental matrix (rank 2, unit Frobenius norm up to
the final refinement)
and it explains the ground-truth inliers.  Parity of the device path with
this oracle is tested in test_gpu_fundamental.py.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r'''
import sys
sys.path.insert(0, %r); sys.path.insert(0, %r)
import numpy as np, oracle_py
from pygcransac import synthetic
sc = synthetic.fundamental_scene(synthetic.scene_seed(7, %d), n=1500, inlier_ratio=0.5, plane_fraction=%r)
p = oracle_py.params(1.0, 0.99, 0.975, 100000, 50, 0, 50, 8.0, %d)
rc, F, m, st = oracle_py.find_fundamental(sc["corr"], 480, 640, 480, 640, p)
np.save(sys.argv[1], np.concatenate([[rc, (m & sc["gt"]).sum(), sc["gt"].sum()], F.ravel()]))
'''


@pytest.mark.parametrize("seed,plane", [(0, 0.8), (2, 0.9)])
def test_degensac_fires_and_returns_a_fundamental_matrix(tmp_path, seed, plane):
    out = str(tmp_path / "r.npy")
    code = SCRIPT % (os.path.join(ROOT, "graph-cut-ransac_amd"), os.path.join(ROOT, "oracle"), seed, plane, seed)
    env = dict(os.environ, ORACLE_TRACE_DEGENSAC="1")
    res = subprocess.run([sys.executable, "-c", code, out], env=env, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    assert "degensac: degenerate sample" in res.stderr
    r = np.load(out)
    rc, tp, gt = int(r[0]), int(r[1]), int(r[2])
    F = r[3:].reshape(3, 3)
    assert rc > 0
    s = np.linalg.svd(F, compute_uv=False)
    assert s[2] < 1e-8 * s[0]
    assert abs(np.linalg.norm(F) - 1.0) < 1e-6
    assert tp >= 0.7 * gt
