"""Learned guidance on the device against the CPU oracle (oracle_mlp.hpp).

Bars:
* features: bit-identical to the oracle (float arithmetic in the same order,
  neighbours by (squared distance, index));
* forward: logits within 2e-2 + 1e-2 |z| of the oracle's BF16-rounding
  forward.  The two differ only in the FP32 summation order inside the
  matrix unit and in tanh/exp ulps, which can move an activation across a
  BF16 rounding boundary (one BF16 ulp, 2^-8 relative);
* gradient (one device forward + backward over 1,000 rows, two weight-
  gradient chunks with padding): per parameter block, relative L2 error
  below 2e-2 and cosine above 0.999 against the oracle's backward pass with
  the same rounding points; loss within 1e-4 relative;
* training: the device's hash-sampled mini-batch, gradient and Adam step
  equal the gradient entry point on the same rows followed by the oracle's
  Adam step (to 1e-6 relative);
* guided batch: identical to gcr_find_batch with PROSAC on the rows sorted
  (stable, descending) by the device's logits, and to the oracle's PROSAC
  run on those rows (masks and statistics identical, models to 1e-12);
* a network trained here (1,024 scenes labelled by plain GC-RANSAC) and the
  committed network (pygcransac/data/guided_mlp.npz, 50,000 scenes) find
  the true inliers of held-out 5 %-inlier pairs that the uniform sampler
  misses at the same iteration budget.
"""
import numpy as np
import pytest

from pygcransac import synthetic

pytestmark = pytest.mark.gpu

OFF = dict(W1=0, b1=1024, W2=1152, b2=17536, W3=17664, b3=34048, w4=34176, b4=34304)
SIZES = dict(W1=1024, b1=128, W2=16384, b2=128, W3=16384, b3=128, w4=128, b4=1)


def _net(seed):
    from pygcransac import guided
    p = guided.init_params(seed)
    rng = np.random.default_rng(seed + 1)
    for b in ("b1", "b2", "b3"):
        p[OFF[b]:OFF[b] + 128] = rng.normal(size=128) * 0.1
    p[OFF["b4"]] = -0.3
    return p


def _scene_pair(seed, n, eps, kind="fundamental"):
    if kind == "fundamental":
        sc = synthetic.fundamental_scene(seed, n=n, inlier_ratio=eps)
    else:
        sc = synthetic.homography_scene(seed, n=n, inlier_ratio=eps)
    return dict(correspondences=sc["corr"], ratio=synthetic.descriptor_ratio(seed, sc["gt"]), h1=480, w1=640,
                h2=480, w2=640), sc


def test_features_bit_exact(engine, oracle):
    from pygcransac import guided
    pairs = []
    for i, (n, w, h) in enumerate([(1, 640, 480), (5, 640, 480), (300, 1024, 768), (700, 640, 480),
                                   (257, 800, 600)]):
        rng = np.random.default_rng(40 + i)
        corr = np.c_[rng.uniform(0, w, (n, 1)), rng.uniform(0, h, (n, 1)), rng.uniform(0, w, (n, 1)),
                     rng.uniform(0, h, (n, 1))]
        if n > 10:
            corr[7] = corr[3]  # a duplicated point: distance ties
        pairs.append(dict(correspondences=corr, ratio=rng.uniform(size=n), h1=h, w1=w, h2=h, w2=w))
    got = guided.features(pairs)
    for pb, f in zip(pairs, got):
        want = oracle.mlp_features(pb["correspondences"], pb["h1"], pb["w1"], pb["h2"], pb["w2"], pb["ratio"])
        assert np.array_equal(f.view(np.uint32), want.view(np.uint32)), len(f)


def test_features_bit_exact_clustered_and_outside_the_image(engine, oracle):
    """The grid neighbour search against the oracle's full scan where it is
    hardest: dense clusters (searches that stop at ring 0 or 1), a block of 30
    identical points (ties broken by index), sparse points far outside the
    image (clamped into the edge cells, searches over many rings), a NaN row
    and pairs of every size from 2 to 9 (fewer than 8 neighbours)."""
    from pygcransac import guided
    rng = np.random.default_rng(77)
    pairs = []
    w, h = 640, 480
    cl = [rng.normal(loc=(rng.uniform(0, w), rng.uniform(0, h)), scale=3.0, size=(400, 2)) for _ in range(5)]
    x1 = np.concatenate(cl + [np.tile([[100.0, 200.0]], (30, 1)), rng.uniform(-3000, 4000, (40, 2)),
                              rng.uniform(0, w, (500, 2))])
    corr = np.c_[x1, x1 + rng.normal(scale=5.0, size=x1.shape)]
    corr[17] = np.nan
    pairs.append(dict(correspondences=corr, ratio=rng.uniform(size=len(corr)), h1=h, w1=w, h2=h, w2=w))
    for n in range(2, 10):
        c = np.c_[rng.uniform(0, w, (n, 1)), rng.uniform(0, h, (n, 1)), rng.uniform(0, w, (n, 1)),
                  rng.uniform(0, h, (n, 1))]
        pairs.append(dict(correspondences=c, ratio=rng.uniform(size=n), h1=h, w1=w, h2=h, w2=w))
    got = guided.features(pairs)
    for pb, f in zip(pairs, got):
        want = oracle.mlp_features(pb["correspondences"], pb["h1"], pb["w1"], pb["h2"], pb["w2"], pb["ratio"])
        assert np.array_equal(f.view(np.uint32), want.view(np.uint32)), len(f)


def test_forward_matches_oracle(engine, oracle):
    from pygcransac import guided
    pb, _ = _scene_pair(synthetic.scene_seed(4, 900), 2000, 0.3)
    feat = guided.features([pb])[0]
    for seed in (1, 2):
        p = _net(seed)
        logit, prob = guided.predict(p, feat)
        want = oracle.mlp_logits(p, feat)
        err = np.abs(logit - want)
        print("forward max |dz| = %.3g (max |z| %.3g)" % (err.max(), np.abs(want).max()))
        assert (err <= 2e-2 + 1e-2 * np.abs(want)).all()
        assert np.allclose(prob, 1 / (1 + np.exp(-logit.astype(np.float64))), rtol=1e-5, atol=1e-6)


def test_gradient_matches_oracle(engine, oracle):
    from pygcransac import guided
    pb, sc = _scene_pair(synthetic.scene_seed(4, 901), 1000, 0.3)
    feat = guided.features([pb])[0]
    labels = sc["gt"].astype(np.float32)
    p = _net(3)
    loss, grad = guided.gradient(p, feat, labels)
    lo, go = oracle.mlp_loss_grad(p, feat, labels)
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    for b, off in OFF.items():
        g, w = grad[off:off + SIZES[b]].astype(np.float64), go[off:off + SIZES[b]]
        if b == "W1":
            g, w = g.reshape(128, 8)[:, :7].ravel(), w.reshape(128, 8)[:, :7].ravel()
        rel = np.linalg.norm(g - w) / max(np.linalg.norm(w), 1e-30)
        cos = g @ w / max(np.linalg.norm(g) * np.linalg.norm(w), 1e-30)
        print("%s: rel %.3g cos %.8f" % (b, rel, cos))
        assert rel < 2e-2 and (cos > 0.999 or SIZES[b] == 1), b
    assert (grad.reshape(-1)[OFF["W1"] + 7:OFF["b1"]:8] == 0).all()


def _mix64(x):
    m = (1 << 64) - 1
    x = (x + 0x9E3779B97F4A7C15) & m
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & m
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & m
    return x ^ (x >> 31)


def _batch_rows(seed, step, batch, n):
    m = (1 << 64) - 1
    return np.array([_mix64(seed ^ _mix64((step * 0x100000001B3 + c) & m)) % n for c in range(batch)])


def test_train_step_matches_gradient_and_oracle_adam(engine, oracle):
    from pygcransac import guided
    pb, sc = _scene_pair(synthetic.scene_seed(4, 902), 1500, 0.3)
    feat = guided.features([pb])[0]
    labels = sc["gt"].astype(np.float32)
    p0 = _net(4)
    tr = guided.Trainer(p0)
    seed, batch, lr = 77, 700, 1e-3
    losses = tr.train(feat, labels, steps=1, batch=batch, lr=lr, seed=seed)
    rows = _batch_rows(seed, 1, batch, len(feat))
    assert len(np.unique(rows)) > batch // 2
    loss, g = guided.gradient(p0, feat[rows], labels[rows])
    assert losses[0] == np.float32(loss)
    assert np.array_equal(tr.m, (np.float32(0.1) * g).astype(np.float32)) or np.allclose(tr.m, 0.1 * g, rtol=1e-6)
    want_p, want_m, want_v = oracle.mlp_adam(p0, np.zeros_like(p0), np.zeros_like(p0), g, 1, lr)
    assert np.allclose(tr.params, want_p, rtol=1e-6, atol=1e-7)
    assert np.allclose(tr.v, want_v, rtol=1e-6, atol=1e-20)


def test_training_reduces_loss(engine):
    from pygcransac import guided
    pb, sc = _scene_pair(synthetic.scene_seed(4, 903), 4000, 0.3)
    feat = guided.features([pb])[0]
    labels = sc["gt"].astype(np.float32)
    tr = guided.Trainer(seed=5)
    losses = tr.train(feat, labels, steps=150, batch=1024, lr=2e-3, seed=1)
    print("loss %.4f -> %.4f" % (losses[:10].mean(), losses[-10:].mean()))
    assert losses[-10:].mean() < 0.7 * losses[:10].mean()
    logit, _ = guided.predict(tr.params, feat)
    # ranking quality: the top tenth by logit is mostly true inliers
    top = np.argsort(-logit, kind="stable")[: len(logit) // 10]
    assert sc["gt"][top].mean() > 0.8


def _sorted_problem(pb, logit):
    order = np.argsort(-logit, kind="stable")
    return order, pb["correspondences"][order]


def test_guided_batch_equals_prosac_on_sorted_rows(engine, oracle):
    import pygcransac
    from pygcransac import guided
    p = _net(6)
    probs = []
    for i, (kind, n, eps) in enumerate([("fundamental", 1500, 0.4), ("homography", 1200, 0.5),
                                       ("fundamental", 800, 0.5)]):
        pb, _ = _scene_pair(synthetic.scene_seed(4, 910 + i), n, eps, kind)
        pb.update(kind=kind, threshold=(3.0 if kind == "homography" else 1.0), conf=0.99,
                  spatial_coherence_weight=0.975, max_iters=5000, min_iters=50, use_sprt=False, neighborhood=0,
                  neighborhood_size=8.0, lo_number=50, seed=300 + i)
        probs.append(pb)
    res, logits = guided.find_batch_guided(probs, p, return_logits=True)
    plain = []
    orders = []
    for pb, lg in zip(probs, logits):
        order, rows = _sorted_problem(pb, lg)
        orders.append(order)
        q = dict(pb)
        q["correspondences"] = rows
        q["sampler"] = 1
        del q["ratio"]
        plain.append(q)
    ref = pygcransac.findBatch(plain)
    for (m, mask), (mr, maskr), order in zip(res, ref, orders):
        unsorted = np.zeros_like(maskr)
        unsorted[order] = maskr
        assert np.array_equal(mask, unsorted)
        assert (m is None) == (mr is None)
        if m is not None:
            assert np.array_equal(m, mr)
    # the oracle's PROSAC run on the same sorted rows (first two pairs)
    for i in range(2):
        q = plain[i]
        pr = oracle.params(q["threshold"], 0.99, 0.975, 5000, 50, 1, 50, 8.0, q["seed"])
        if q["kind"] == "homography":
            rc, Mo, mo, so = oracle.find_homography(q["correspondences"], 480, 640, 480, 640, pr)
        else:
            rc, Mo, mo, so = oracle.find_fundamental(q["correspondences"], 480, 640, 480, 640, pr)
        assert np.array_equal(ref[i][1], mo.astype(bool))
        assert guided.last_stats[i]["iteration_number"] == so["iteration_number"]
        assert np.allclose(ref[i][0], Mo, rtol=1e-12, atol=1e-12)


@pytest.mark.timeout(300)
def test_trained_network_beats_uniform_sampler_at_5_percent(engine):
    from pygcransac import train_guided
    # training data: plain GC-RANSAC masks on 1,024 30 %-inlier fundamental
    # scenes of 5,000 correspondences (device labels, features and training)
    feat, labels = train_guided.build_training_set(1024, workers=0, log=lambda s: None)
    assert 0.25 < labels.mean() < 0.35
    tr, losses = train_guided.train(feat, labels, steps=3000, batch=8192, lr=1e-3, seed=0, log=lambda s: None)
    assert losses[-200:].mean() < 0.7 * losses[:20].mean()
    # held-out 5 %-inlier pairs of 4,000 correspondences, equal iteration caps
    probs, gts = train_guided.held_out(24, max_iters=2000, first=5000)
    g, u = train_guided.evaluate(tr.params, probs, gts)
    print("true-inlier recall: guided %s, uniform %s" % (np.round(g, 3), np.round(u, 3)))
    assert g.mean() > 0.9 and (g < 0.5).sum() <= 1
    assert g.mean() > u.mean() + 0.5


def test_committed_network_on_held_out_pairs(engine):
    from pygcransac import guided, train_guided
    params = guided.load()
    probs, gts = train_guided.held_out(32, max_iters=10000, first=0)
    g, u = train_guided.evaluate(params, probs, gts)
    print("committed network, true-inlier recall: guided %.4f (min %.3f), uniform %.4f" % (g.mean(), g.min(),
                                                                                          u.mean()))
    assert g.mean() > 0.95 and (g < 0.5).sum() == 0
    assert u.mean() < 0.3
