"""The C-ABI library loads (no GPU needed) and exports every entry point
include/gcr.h declares; the Python binding's export list matches the header.
No compute call is made here."""
import ctypes
import os
import re

from pygcransac import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "gcr.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gcr_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    for must in ("gcr_find_homography", "gcr_find_fundamental_matrix", "gcr_find_essential_matrix",
                 "gcr_find_batch", "gcr_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_native.library_path())
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_export_list_matches_header():
    assert sorted(_native.EXPORTS) == _declared()


# ---- batch entry points refuse bad pairs before any device work ------------

def _batch_args(kinds, rows=20, K=None, tweak=None):
    import numpy as np
    n = len(kinds)
    kinds = np.asarray(kinds, np.int32)
    offsets = np.arange(n + 1, dtype=np.int64) * rows
    corr = np.random.default_rng(0).uniform(0, 100, (n * rows, 4))
    sizes = np.full(4 * n, 100, np.int32)
    params = (_native.GcrParams * n)()
    L = _native.lib()
    for i in range(n):
        L.gcr_default_params(ctypes.byref(params[i]), int(kinds[i]) if 0 <= kinds[i] <= 2 else 0)
        params[i].neighborhood = 0
        params[i].sampler = 0
        if tweak:
            tweak(i, params[i])
    return dict(n=n, kinds=kinds, offsets=offsets, corr=corr, sizes=sizes, K=K, params=params,
                models=np.zeros(9 * n), masks=np.zeros(n * rows, np.uint8), ninl=np.zeros(n, np.int32),
                ratios=np.zeros(n * rows), mlp=np.zeros(_native.lib().gcr_mlp_num_params(), np.float32))


def _vp(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _call_batch(a):
    return _native.lib().gcr_find_batch(a["n"], _vp(a["kinds"]), _vp(a["offsets"]), _vp(a["corr"]), 0,
                                        _vp(a["sizes"]), _vp(a["K"]), a["params"], _vp(a["models"]),
                                        _vp(a["masks"]), _vp(a["ninl"]), None)


def _call_guided(a):
    return _native.lib().gcr_find_batch_guided(a["n"], _vp(a["kinds"]), _vp(a["offsets"]), _vp(a["corr"]), 0,
                                               _vp(a["sizes"]), _vp(a["K"]), _vp(a["ratios"]), _vp(a["mlp"]),
                                               a["params"], _vp(a["models"]), _vp(a["masks"]), _vp(a["ninl"]),
                                               None, None)


def _last_error():
    L = _native.lib()
    L.gcr_last_error.restype = ctypes.c_char_p
    return L.gcr_last_error().decode()


def _bad_cases():
    import numpy as np
    good_K = np.tile(np.r_[np.eye(3).ravel(), np.eye(3).ravel()], 2)
    sing_K = good_K.copy()
    sing_K[18 + 8] = 0.0  # pair 1: K1 singular
    sing_K[18 + 4] = 0.0

    def conf(i, p):
        if i == 1:
            p.confidence = 1.0

    def thr(i, p):
        if i == 1:
            p.threshold = -1.0

    def solver(i, p):
        if i == 1:
            p.solver = 1

    def iters(i, p):
        if i == 1:
            p.max_iters = -5

    def nbh(i, p):
        if i == 1:
            p.neighborhood_size = 0.5
            p.spatial_coherence_weight = 0.5

    def nb_id(i, p):
        if i == 1:
            p.neighborhood = 5
            p.spatial_coherence_weight = 0.5

    def radius_inf(i, p):
        if i == 1:
            p.neighborhood = 1
            p.neighborhood_size = float("inf")
            p.spatial_coherence_weight = 0.5

    return [
        ("kind 3 (internal estimator id)", dict(kinds=[0, 3]), -1, "unknown problem kind"),
        ("neighbourhood id 5", dict(kinds=[0, 1], tweak=nb_id), -1, "neighborhood-graph identifier"),
        ("infinite radius", dict(kinds=[0, 1], tweak=radius_inf), -1, "radius"),
        ("kind -1", dict(kinds=[0, -1]), -1, "unknown problem kind"),
        ("essential without K", dict(kinds=[0, 2]), -1, "K1 and K2"),
        ("essential with singular K", dict(kinds=[2, 2], K=sing_K), -1, "invertible"),
        ("confidence 1", dict(kinds=[0, 1], tweak=conf), -1, "conf"),
        ("negative threshold", dict(kinds=[0, 1], tweak=thr), -1, "threshold"),
        ("solver 1", dict(kinds=[0, 1], tweak=solver), -2, "solver"),
        ("negative max_iters", dict(kinds=[0, 1], tweak=iters), -1, "max_iters"),
        ("neighborhood_size 0.5", dict(kinds=[0, 1], tweak=nbh), -1, "neighborhood_size"),
        ("too few rows", dict(kinds=[0, 1], rows=5), -1, "n>=7"),
    ]


def test_batch_entry_points_refuse_bad_pairs_before_device_work():
    """Both batch entry points apply the single-pair rules to every pair and
    refuse bad ones with the pair's index, before any device call (this
    container has no GPU, so reaching the device would return -3)."""
    for name, kw, code, msg in _bad_cases():
        for call in (_call_batch, _call_guided):
            a = _batch_args(**kw)
            rc = call(a)
            assert rc == code, (name, call.__name__, rc, _last_error())
            err = _last_error()
            assert msg in err and "pair 1" in err, (name, err)


def test_batch_good_pairs_reach_the_device_check():
    """Valid pairs pass the host checks; without a GPU the call then stops
    with GCR_ERR_DEVICE (-3)."""
    import numpy as np
    L = _native.lib()
    if L.gcr_device_count() > 0:
        import pytest
        pytest.skip("a GPU is visible: the call would run")
    K = np.tile(np.r_[np.eye(3).ravel(), np.eye(3).ravel()], 3)
    a = _batch_args([0, 1, 2], rows=20, K=K)
    assert _call_batch(a) == -3


def test_unknown_neighbourhood_returns_no_model_before_device_work():
    """The drivers print a message and return 0 inliers for an unknown
    neighbourhood id (py/gcransac_python.cpp:1681-1686, 2604-2609,
    4302-4307); the single-pair entry points return no model before any
    device call.  A homography with space partitioning on switches to the
    grid (:4282-4286) and so reaches the device check."""
    import numpy as np
    import pygcransac
    corr = np.random.default_rng(1).uniform(0, 100, (50, 4))
    H, mask = pygcransac.findHomography(corr, 100, 100, 100, 100, None, neighborhood=7,
                                        use_space_partitioning=False, use_sprt=False, sampler=0, seed=1)
    assert H is None and mask.shape == (50,) and not mask.any()
    assert "neighborhood-graph identifier" in _last_error()
    F, mask = pygcransac.findFundamentalMatrix(corr, 100, 100, 100, 100, None, spatial_coherence_weight=0.2,
                                               neighborhood=7, use_sprt=False, sampler=0, seed=1)
    assert F is None and not mask.any()
    L = _native.lib()
    if L.gcr_device_count() <= 0:
        import pytest
        with pytest.raises(RuntimeError):
            pygcransac.findHomography(corr, 100, 100, 100, 100, None, neighborhood=7,
                                      use_space_partitioning=True, use_sprt=False, sampler=0, seed=1)
