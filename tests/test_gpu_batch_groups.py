"""A batch split into concurrent groups (one engine, stream and host pool per
group of consecutive pairs) gives the results of the same batch run as one
group, bit for bit: models, masks and statistics per pair, with the
correspondences given in host memory and in device memory.
"""
import ctypes

import numpy as np
import pytest

from pygcransac import _native, hipmem, synthetic

pytestmark = pytest.mark.gpu


def _problems():
    out = []
    for i in range(24):
        n = 300 + 53 * i
        if i % 2 == 0:
            sc = synthetic.homography_scene(synthetic.scene_seed(9, i), n=n, inlier_ratio=0.5)
            kind, thr = "homography", 3.0
        else:
            sc = synthetic.fundamental_scene(synthetic.scene_seed(9, i), n=n, inlier_ratio=0.55)
            kind, thr = "fundamental", 1.0
        out.append(dict(kind=kind, correspondences=sc["corr"], h1=480, w1=640, h2=480, w2=640, threshold=thr,
                        conf=0.99, spatial_coherence_weight=0.975, max_iters=100000, min_iters=50, use_sprt=False,
                        sampler=0, neighborhood=0, neighborhood_size=8.0, lo_number=50, seed=300 + i))
    return out


def _run(monkeypatch, probs, groups):
    import pygcransac
    monkeypatch.setenv("GCR_BATCH_GROUPS", str(groups))
    res = pygcransac.findBatch(probs)
    return res, pygcransac.last_batch_stats()


def _same(a, b, sa, sb):
    for i, ((Ma, ma), (Mb, mb)) in enumerate(zip(a, b)):
        assert np.array_equal(ma, mb), i
        assert (Ma is None) == (Mb is None), i
        if Ma is not None:
            assert np.array_equal(Ma.view(np.uint64), Mb.view(np.uint64)), i
        for key in ("iteration_number", "local_optimization_number", "graph_cut_number", "num_inliers"):
            assert sa[i][key] == sb[i][key], (i, key)


@pytest.mark.parametrize("groups", [3, 5])
def test_grouped_batch_identical_to_one_group(monkeypatch, groups):
    probs = _problems()
    one, s1 = _run(monkeypatch, probs, 1)
    assert sum(m.sum() for _, m in one) > 5000
    grouped, sg = _run(monkeypatch, probs, groups)
    _same(one, grouped, s1, sg)


def test_grouped_batch_device_input(monkeypatch):
    # device input: each group's rows start after the earlier groups' rows
    probs = _problems()
    one, _ = _run(monkeypatch, probs, 1)
    monkeypatch.setenv("GCR_BATCH_GROUPS", "4")
    import pygcransac
    n = len(probs)
    corr = np.ascontiguousarray(np.concatenate([p["correspondences"] for p in probs]))
    offsets = np.zeros(n + 1, np.int64)
    offsets[1:] = np.cumsum([len(p["correspondences"]) for p in probs])
    kinds = np.array([0 if p["kind"] == "homography" else 1 for p in probs], np.int32)
    sizes = np.tile(np.array([480, 640, 480, 640], np.int32), n)
    params = (_native.GcrParams * n)()
    for i, p in enumerate(probs):
        params[i] = pygcransac._params(int(kinds[i]), threshold=p["threshold"], conf=0.99,
                                       spatial_coherence_weight=0.975, max_iters=100000, min_iters=50,
                                       use_sprt=False, min_inlier_ratio_for_sprt=1e-5, sampler=0, neighborhood=0,
                                       neighborhood_size=8.0, use_space_partitioning=False, lo_number=50,
                                       sampler_variance=0.1, solver=0, seed=p["seed"])
    dev = hipmem.DeviceArray(corr)
    try:
        models = np.zeros(9 * n)
        masks = np.zeros(int(offsets[-1]), np.uint8)
        ninl = np.zeros(n, np.int32)
        stats = (_native.GcrStats * n)()
        vp = ctypes.c_void_p
        rc = _native.lib().gcr_find_batch(
            n, kinds.ctypes.data_as(vp), offsets.ctypes.data_as(vp), ctypes.c_void_p(dev.ptr.value), 1,
            sizes.ctypes.data_as(vp), None, params, models.ctypes.data_as(vp), masks.ctypes.data_as(vp),
            ninl.ctypes.data_as(vp), stats)
        _native.raise_for(rc)
    finally:
        dev.free()
    for i, (M, m) in enumerate(one):
        assert np.array_equal(masks[offsets[i]:offsets[i + 1]].astype(bool), m), i
        if M is not None:
            assert np.array_equal(models[9 * i:9 * i + 9].view(np.uint64), M.ravel().view(np.uint64)), i
