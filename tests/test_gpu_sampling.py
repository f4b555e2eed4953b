"""The samplers' random draws on the device (k_sample.hip) against the
oracle's samplers: the uniform sampler (inc/samplers/uniform_sampler.h:97-115)
and PROSAC (inc/samplers/prosac_sampler.h:176-237), both through
inc/uniform_random_generator.h:43-125.  The same indices, chunk after chunk
of one stream (each chunk its own launch, so the generator position stored
between launches is checked too).

Uniform cases: a pool of the usual size, tiny pools where repeated indices are
redrawn often (pool == m: every sample is a permutation), and ranges where the
bounded-integer step rejects many words (2^28 + 5: 6 %; 1.5e9: 30 %; those run
one sample per batch).  PROSAC cases: the growing subsets of the first
thousands of samples, and a small point set run past the convergence length
(100,000 samples), where every index of a sample is drawn from all points."""
import numpy as np
import pytest

from pygcransac import _native

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,pool,m,chunks", [
    (0, 2000, 4, [256]),
    (12345, 5000, 7, [256, 256, 512, 1024, 3000]),
    (7, 10, 5, [100, 37, 1, 0, 64]),
    (99, 7, 7, [50, 50]),
    (3, 2 ** 28 + 5, 7, [70, 1, 300]),
    (17, 1_500_000_000, 7, [130, 2, 257]),
    (5, 2 ** 31 - 1, 4, [64, 65]),
])
def test_device_uniform_draws_equal_the_oracle(engine, oracle, seed, pool, m, chunks):
    got = _native.sampler_draws(0, seed, pool, m, chunks)
    want = oracle.uniform_draws(seed, pool, m, int(sum(chunks)))
    assert np.array_equal(got.astype(np.int64), want)


def test_pool_smaller_than_the_sample_gives_minus_one(engine):
    got = _native.sampler_draws(0, 4, 3, 4, [5, 2])
    assert got.shape == (7, 4) and (got == -1).all()


@pytest.mark.parametrize("seed,n,m,chunks", [
    (11, 4000, 7, [256, 256, 512, 1024, 2048, 4096]),
    (2, 2000, 4, [300, 1, 700]),
    (8, 30, 7, [60000, 40100, 300]),
])
def test_device_prosac_draws_equal_the_oracle(engine, oracle, seed, n, m, chunks):
    got = _native.sampler_draws(1, seed, n, m, chunks)
    want = oracle.prosac_draws(seed, n, m, int(sum(chunks)))
    assert np.array_equal(got.astype(np.int64), want)


def _batch(sampler):
    from pygcransac import synthetic
    out = []
    for i in range(10):
        n = 400 + 97 * i
        if i % 2 == 0:
            sc = synthetic.homography_scene(synthetic.scene_seed(21, i), n=n, inlier_ratio=0.4)
            kind, thr = "homography", 3.0
        else:
            sc = synthetic.fundamental_scene(synthetic.scene_seed(21, i), n=n, inlier_ratio=0.45)
            kind, thr = "fundamental", 1.0
        out.append(dict(kind=kind, correspondences=sc["corr"], h1=480, w1=640, h2=480, w2=640, threshold=thr,
                        conf=0.99, spatial_coherence_weight=0.975, max_iters=100000, min_iters=50, use_sprt=False,
                        sampler=sampler, neighborhood=0, neighborhood_size=8.0, lo_number=50, seed=500 + i))
    return out


@pytest.mark.parametrize("sampler,env", [(0, "GCR_DEVICE_SAMPLING"), (1, "GCR_DEVICE_PROSAC")])
def test_full_runs_identical_with_device_and_host_draws(monkeypatch, engine, sampler, env):
    """Whole GC-RANSAC runs (homography and fundamental pairs, DEGENSAC
    included) give the same bits with the sampler's draws made on the device
    and on the host.  The host draws the next chunk's samples ahead while the
    device works and a chunk takes them first, so a sample taken out of
    stream order, twice or not at all changes the host runs."""
    import pygcransac
    probs = _batch(sampler)
    monkeypatch.setenv("GCR_DEVICE_SAMPLING", "0")
    monkeypatch.setenv("GCR_DEVICE_PROSAC", "0")
    host = pygcransac.findBatch(probs)
    sh = pygcransac.last_batch_stats()
    monkeypatch.setenv("GCR_DEVICE_SAMPLING", "1")
    monkeypatch.setenv(env, "1")
    dev = pygcransac.findBatch(probs)
    sd = pygcransac.last_batch_stats()
    assert sum(m.sum() for _, m in host) > 1000
    for i, ((Ma, ma), (Mb, mb)) in enumerate(zip(host, dev)):
        assert np.array_equal(ma, mb), i
        assert (Ma is None) == (Mb is None), i
        if Ma is not None:
            assert np.array_equal(Ma.view(np.uint64), Mb.view(np.uint64)), i
        for key in ("iteration_number", "local_optimization_number", "graph_cut_number", "num_inliers"):
            assert sh[i][key] == sd[i][key], (i, key)
