"""gcr_score_models_noalloc (device buffers in and out, the library's fixed
device scratch, no allocation) against gcr_score_models on the same inputs:
identical counts, values and masks, on two calls in a row, the second with
fewer points and fewer models than the first."""
import ctypes

import numpy as np
import pytest

from pygcransac import _native, hipmem, synthetic

pytestmark = pytest.mark.gpu


def _models(kind, sc, rng, k):
    base = sc["H"] if kind == 0 else sc["F"]
    out = []
    for j in range(k):
        M = base * (1.0 + rng.normal(0, 1e-3 * (j % 5), (3, 3)))
        out.append(M.ravel())
    return np.ascontiguousarray(np.array(out))


def _noalloc(kind, pts, models, sq, with_masks):
    n, k = pts.shape[0], models.shape[0]
    d_pts = hipmem.DeviceArray(pts)
    d_mod = hipmem.DeviceArray(models)
    d_cnt = hipmem.DeviceArray(np.full(k, -7, np.int64))
    d_val = hipmem.DeviceArray(np.full(k, np.nan))
    d_msk = hipmem.DeviceArray(np.full(k * n, 9, np.uint8)) if with_masks else None
    try:
        rc = _native.lib().gcr_score_models_noalloc(kind, d_pts.ptr, n, d_mod.ptr, k, ctypes.c_double(sq),
                                                    d_cnt.ptr, d_val.ptr, d_msk.ptr if d_msk else None, None)
        _native.raise_for(rc)
        assert hipmem.hip().hipDeviceSynchronize() == 0
        cnt = d_cnt.to_host(np.int64, k)
        val = d_val.to_host(np.float64, k)
        msk = d_msk.to_host(np.uint8, k * n).reshape(k, n) if d_msk else None
    finally:
        for a in (d_pts, d_mod, d_cnt, d_val, d_msk):
            if a is not None:
                a.free()
    return cnt, val, msk


@pytest.mark.parametrize("kind", [0, 1])
def test_noalloc_scoring_matches_allocating_entry_point(kind):
    rng = np.random.default_rng(100 + kind)
    if kind == 0:
        sc = synthetic.homography_scene(synthetic.scene_seed(0, 7))
        sq = (1.5 * 3.0) ** 2
    else:
        sc = synthetic.fundamental_scene(synthetic.scene_seed(1, 7))
        sq = (1.5 * 1.0) ** 2
    pts = np.ascontiguousarray(sc["corr"])
    # first call: all points, 700 models; second: fewer points and models
    for n, k, masks in ((pts.shape[0], 700, True), (pts.shape[0] // 3 + 17, 130, True), (501, 33, False)):
        p = np.ascontiguousarray(pts[:n])
        M = _models(kind, sc, rng, k)
        print("noalloc kind=%d n=%d k=%d masks=%d" % (kind, n, k, masks), flush=True)
        r0 = _native.score_models(kind, p, M, sq, want_masks=masks)
        c0, v0, m0 = r0 if masks else (r0[0], r0[1], None)
        c1, v1, m1 = _noalloc(kind, p, M, sq, masks)
        assert np.array_equal(c0.astype(np.int64), c1)
        assert np.array_equal(np.asarray(v0).view(np.uint64), v1.view(np.uint64))
        if masks:
            assert np.array_equal(np.asarray(m0, np.uint8), m1)
        assert c1.max() > 0


def test_noalloc_refuses_too_many_models():
    L = _native.lib()
    big = 65537
    rc = L.gcr_score_models_noalloc(0, ctypes.c_void_p(16), 10, ctypes.c_void_p(16), big, ctypes.c_double(1.0),
                                    ctypes.c_void_p(16), ctypes.c_void_p(16), None, None)
    assert rc == _native.GCR_ERR_INVALID_ARGUMENT
    assert b"65536" in L.gcr_last_error()
