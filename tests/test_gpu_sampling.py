"""The uniform sampler's draws on the device (k_sample.hip) against the
oracle's sampler (inc/samplers/uniform_sampler.h:97-115 through
inc/uniform_random_generator.h:43-125): the same indices, chunk after chunk
of one stream (each chunk its own launch, so the generator position stored
between launches is checked too).

The cases cover a pool of the usual size, tiny pools where repeated indices
are redrawn often (pool == m: every sample is a permutation), and ranges
where the bounded-integer step rejects many words (2^28 + 5: 6 %; 1.5e9:
30 %; those run one sample per batch and give back read-ahead words across
block boundaries)."""
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
    got = _native.uniform_samples(seed, pool, m, chunks)
    want = oracle.uniform_draws(seed, pool, m, int(sum(chunks)))
    assert np.array_equal(got.astype(np.int64), want)


def test_pool_smaller_than_the_sample_gives_minus_one(engine):
    got = _native.uniform_samples(4, 3, 4, [5, 2])
    assert got.shape == (7, 4) and (got == -1).all()
