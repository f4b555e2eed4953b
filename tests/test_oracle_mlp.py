"""Pins the learned-guidance oracle (oracle/oracle_mlp.hpp) on the CPU.

The oracle is the checker of the device's feature, forward, backward and
Adam kernels, so each of its parts is checked against something it does not
restate itself: scipy's k-d tree for the neighbour sets, central finite
differences of the loss for the backward pass, closed forms for GELU,
softplus/BCE and the first Adam step, and numpy's float32 bits for BF16
rounding.  A planted mistake (a dropped GELU-derivative term, a transposed
weight, a neighbour off by one) fails one of these.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "graph-cut-ransac_amd"))

import oracle_py  # noqa: E402
from pygcransac import synthetic  # noqa: E402

OFF = dict(W1=0, b1=1024, W2=1152, b2=17536, W3=17664, b3=34048, w4=34176, b4=34304)
NP = 34305


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NP, np.float32)
    p[OFF["W1"]:OFF["b1"]] = (rng.normal(size=1024) * np.sqrt(2 / 7) * scale).astype(np.float32)
    p[OFF["W1"] + 7:OFF["b1"]:8] = 0.0
    p[OFF["b1"]:OFF["W2"]] = rng.normal(size=128) * 0.1
    p[OFF["W2"]:OFF["b2"]] = rng.normal(size=16384) * np.sqrt(2 / 128) * scale
    p[OFF["b2"]:OFF["W3"]] = rng.normal(size=128) * 0.1
    p[OFF["W3"]:OFF["b3"]] = rng.normal(size=16384) * np.sqrt(2 / 128) * scale
    p[OFF["b3"]:OFF["w4"]] = rng.normal(size=128) * 0.1
    p[OFF["w4"]:OFF["b4"]] = rng.normal(size=128) * np.sqrt(1 / 128) * scale
    p[OFF["b4"]] = 0.2
    return p


def _feat(seed, n):
    rng = np.random.default_rng(seed)
    f = rng.uniform(-1, 1, size=(n, 8)).astype(np.float32)
    f[:, 4] = rng.uniform(0, 1, size=n)
    f[:, 7] = 0.0
    return f


def test_num_params():
    assert oracle_py.mlp_num_params() == NP


def test_round_bf16_ties_to_even():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(size=5000) * 10.0 ** rng.integers(-6, 6, 5000),
                        [1.0, -2.5, 0.0, 1.00390625, 1.01171875, 3.0e-39]]).astype(np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    low = u & 0xFFFF
    up = u >> 16
    # ties to even: round up when low > 0x8000, or == 0x8000 with odd upper half
    up = up + ((low > 0x8000) | ((low == 0x8000) & ((up & 1) == 1)))
    want = (up << 16).astype(np.uint32).view(np.float32)
    got = oracle_py.mlp_round_bf16(x)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # exact ties: 1 + 2^-8 is halfway between 1 and 1 + 2^-7 -> even (1.0)
    assert oracle_py.mlp_round_bf16(np.array([1.00390625], np.float32))[0] == 1.0
    assert oracle_py.mlp_round_bf16(np.array([1.01171875], np.float32))[0] == 1.015625


def test_gelu_closed_form():
    z = np.array([-6.0, -1.0, 0.0, 0.5, 1.0, 3.0, 8.0])
    g, dg = oracle_py.mlp_gelu(z)
    c = np.sqrt(2 / np.pi)
    want = 0.5 * z * (1 + np.tanh(c * (z + 0.044715 * z ** 3)))
    assert np.allclose(g, want, rtol=1e-14, atol=1e-15)
    assert g[2] == 0.0 and abs(g[4] - 0.8411919906082768) < 1e-12
    h = 1e-6
    num = (oracle_py.mlp_gelu(z + h)[0] - oracle_py.mlp_gelu(z - h)[0]) / (2 * h)
    assert np.allclose(dg, num, rtol=1e-7, atol=1e-9)
    assert abs(dg[2] - 0.5) < 1e-15


def test_knn_against_kdtree():
    from scipy.spatial import cKDTree
    sc = synthetic.fundamental_scene(synthetic.scene_seed(4, 3), n=700, inlier_ratio=0.2)
    corr = sc["corr"]
    nb = oracle_py.mlp_knn(corr)
    p32 = corr[:, :2].astype(np.float32)
    # brute-force order by (float32 squared distance, index), self excluded
    for q in range(0, len(corr), 7):
        d = ((p32[q] - p32) ** 2).sum(axis=1, dtype=np.float32)
        d[q] = np.inf
        order = np.lexsort((np.arange(len(d)), d))[:8]
        assert np.array_equal(nb[q], order), q
    # and the sets agree with a k-d tree on the float64 points (no ties here)
    tree = cKDTree(corr[:, :2])
    _, idx = tree.query(corr[:, :2], k=9)
    for q in range(len(corr)):
        assert set(nb[q]) == set(idx[q][idx[q] != q][:8]), q


def test_knn_small_and_duplicate_points():
    corr = np.array([[1.0, 1.0, 2.0, 2.0], [1.0, 1.0, 5.0, 1.0], [3.0, 1.0, 3.0, 3.0]])
    nb = oracle_py.mlp_knn(corr)
    assert list(nb[0][:2]) == [1, 2] and (nb[0][2:] == -1).all()
    assert list(nb[1][:2]) == [0, 2]
    f = oracle_py.mlp_features(corr, 480, 640, 480, 640, np.array([0.1, 0.2, 0.3]))
    # point 0: own displacement (1, 1); neighbours 1, 2 displace (4, 0), (0, 2)
    sd = np.float32(2.0) / np.float32(640)
    assert f[0, 5] == np.float32((np.float32(4.0) / np.float32(2) - np.float32(1.0)) * sd)
    assert f[0, 6] == np.float32((np.float32(2.0) / np.float32(2) - np.float32(1.0)) * sd)
    assert np.allclose(f[:, 4], [0.1, 0.2, 0.3]) and (f[:, 7] == 0).all()
    assert f[0, 0] == np.float32(np.float32(1.0) * (np.float32(2.0) / np.float32(640)) - np.float32(1.0))
    one = oracle_py.mlp_features(corr[:1], 480, 640, 480, 640, np.array([0.5]))
    assert one[0, 5] == 0.0 and one[0, 6] == 0.0


def test_features_mean_displacement():
    sc = synthetic.homography_scene(synthetic.scene_seed(4, 1), n=300)
    corr = sc["corr"]
    f = oracle_py.mlp_features(corr, 480, 640, 480, 640, np.zeros(len(corr)))
    nb = oracle_py.mlp_knn(corr)
    c32 = corr.astype(np.float32)
    u = np.stack([c32[:, 2] - c32[:, 0], c32[:, 3] - c32[:, 1]], 1)
    want = (u[nb].mean(axis=1) - u) * (2.0 / 640)
    assert np.allclose(f[:, 5:7], want, rtol=1e-5, atol=1e-6)


def _fd_check(block, idx, p, feat, lab, grad):
    h = 1e-6
    pd = p.astype(np.float64)
    for i in idx:
        a, b = pd.copy(), pd.copy()
        a[i] += h
        b[i] -= h
        la, _ = oracle_py.mlp_loss_grad(p, feat, lab, exact=True, params_d=a, want_grad=False)
        lb, _ = oracle_py.mlp_loss_grad(p, feat, lab, exact=True, params_d=b, want_grad=False)
        num = (la - lb) / (2 * h)
        assert abs(num - grad[i]) <= 1e-6 * max(1.0, abs(num)) + 1e-8, (block, i, num, grad[i])


@pytest.mark.parametrize("block", list(OFF))
def test_backward_matches_finite_differences(block):
    p = _params(7)
    feat = _feat(8, 24)
    lab = (np.random.default_rng(9).uniform(size=24) < 0.4).astype(np.float32)
    _, grad = oracle_py.mlp_loss_grad(p, feat, lab, exact=True)
    lo = OFF[block]
    hi = {"W1": 1024, "b1": 128, "W2": 16384, "b2": 128, "W3": 16384, "b3": 128, "w4": 128, "b4": 1}[block] + lo
    rng = np.random.default_rng(hash(block) & 0xFFFF)
    idx = rng.choice(np.arange(lo, hi), size=min(12, hi - lo), replace=False)
    if block == "W1":
        idx = idx[(idx - lo) % 8 != 7]  # the padded input column carries no gradient
    _fd_check(block, idx, p, feat, lab, grad)


def test_bf16_mode_close_to_exact():
    p = _params(3)
    feat = _feat(4, 64)
    lab = (np.random.default_rng(5).uniform(size=64) < 0.5).astype(np.float32)
    le, ge = oracle_py.mlp_loss_grad(p, feat, lab, exact=True)
    lq, gq = oracle_py.mlp_loss_grad(p, feat, lab, exact=False)
    assert abs(le - lq) < 2e-2 * abs(le)
    cos = ge @ gq / np.linalg.norm(ge) / np.linalg.norm(gq)
    assert cos > 0.99


def test_loss_closed_form():
    # all weights zero: logit = b4 for every point; loss = softplus(b4) - y b4
    p = np.zeros(NP, np.float32)
    p[OFF["b4"]] = 0.75
    feat = _feat(1, 10)
    lab = np.array([1, 0, 1, 1, 0, 0, 0, 1, 0, 1], np.float32)
    loss, grad = oracle_py.mlp_loss_grad(p, feat, lab, exact=True)
    z = 0.75
    want = np.mean(np.log1p(np.exp(z)) - lab * z)
    assert abs(loss - want) < 1e-12
    sig = 1 / (1 + np.exp(-z))
    assert abs(grad[OFF["b4"]] - np.mean(sig - lab)) < 1e-12


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(2)
    p = rng.normal(size=NP).astype(np.float32)
    g = (rng.normal(size=NP) * 10.0 ** rng.integers(-4, 2, NP)).astype(np.float32)
    lr = 1e-3
    p1, m, v = oracle_py.mlp_adam(p, np.zeros(NP), np.zeros(NP), g, 1, lr)
    want = p.astype(np.float64) - lr * g / (np.abs(g.astype(np.float64)) + 1e-8)
    assert np.allclose(p1, want, rtol=1e-6, atol=1e-7)
    one_m_b2 = float(np.float32(1.0) - np.float32(0.999))  # 1 - 0.999 in float32
    assert np.allclose(m, 0.1 * g, rtol=1e-6) and np.allclose(v, one_m_b2 * g.astype(np.float64) ** 2, rtol=1e-5)
