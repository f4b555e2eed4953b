"""pygcransac -- the reference's Python API on the MI355X engine.

Same function names, positional order and keyword defaults as the
reference's pybind11 module (src/pygcransac/src/bindings.cpp:985-1179):

    findHomography(correspondences, h1, w1, h2, w2, probabilities, ...)
    findFundamentalMatrix(correspondences, h1, w1, h2, w2, probabilities, ...)
    findEssentialMatrix(correspondences, K1, K2, h1, w1, h2, w2, probabilities, ...)

Each returns ``(model, mask)``: a 3x3 float64 array (``None`` when no model
was found) and an (N,) bool inlier mask.  Shape errors raise ValueError as
the reference does (bindings.cpp:881-889).  Options the engine does not
implement raise NotImplementedError instead of silently running a different
algorithm.  ``use_sprt`` and ``use_space_partitioning`` are accepted; the
engine scores every hypothesis on all points, which is what those
accelerations approximate.

Extra keyword: ``seed`` (int).  The reference seeds its samplers from
std::random_device; here ``seed=None`` draws a fresh seed per call and an
integer makes the call reproducible.

Everything runs in libgcr_hip.so through its C ABI (include/gcr.h); there is
no host fallback: without the built library or a GPU the calls raise.
"""
import ctypes
import os

import numpy as np

from . import _native
from ._native import GcrParams, GcrStats

__all__ = [
    "findHomography",
    "findFundamentalMatrix",
    "findEssentialMatrix",
    "findBatch",
    "last_stats",
]

_last_stats = None


_last_batch_stats = None


def last_stats():
    """Statistics of the last find* call of this process (dict)."""
    return _last_stats


def last_batch_stats():
    """Per-pair statistics (list of dicts) of the last findBatch call."""
    return _last_batch_stats


def _as_corr(correspondences, min_n, what):
    a = np.ascontiguousarray(np.asarray(correspondences, dtype=np.float64))
    if a.ndim != 2 or a.shape[1] != 4:
        raise ValueError("correspondences should be an array with dims [n,4], n>=%d" % min_n)
    if a.shape[0] < min_n:
        raise ValueError("%s should be an array with dims [n,d], n>=%d" % (what, min_n))
    return a


def _params(problem, threshold, conf, spatial_coherence_weight, max_iters, min_iters, use_sprt,
            min_inlier_ratio_for_sprt, sampler, neighborhood, neighborhood_size, use_space_partitioning,
            lo_number, sampler_variance, solver, seed):
    p = GcrParams()
    _native.lib().gcr_default_params(ctypes.byref(p), problem)
    p.threshold = float(threshold)
    p.confidence = float(conf)
    p.spatial_coherence_weight = float(spatial_coherence_weight)
    p.max_iters = int(max_iters)
    p.min_iters = int(min_iters)
    p.use_sprt = int(bool(use_sprt))
    p.min_inlier_ratio_for_sprt = float(min_inlier_ratio_for_sprt)
    p.sampler = int(sampler)
    p.neighborhood = int(neighborhood)
    p.neighborhood_size = float(neighborhood_size)
    p.use_space_partitioning = int(bool(use_space_partitioning))
    p.lo_number = int(lo_number)
    p.sampler_variance = float(sampler_variance)
    p.solver = int(solver)
    if seed is None:
        seed = int.from_bytes(os.urandom(4), "little")
    p.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return p


def _check_solver(solver):
    if solver != 0:
        raise NotImplementedError("only point correspondences (solver=0) are built")


def _finish(rc, model, mask, stats):
    global _last_stats
    _native.raise_for(rc)
    _last_stats = stats.as_dict()
    m = mask.astype(bool)
    if rc == 0:
        return None, m
    return model.reshape(3, 3).copy(), m


def _probs(probabilities):
    if probabilities is None:
        return None, 0
    pr = np.ascontiguousarray(np.asarray(probabilities, dtype=np.float64).ravel())
    return pr, pr.size


def findHomography(correspondences, h1, w1, h2, w2, probabilities, threshold=1.0, conf=0.99,
                   spatial_coherence_weight=0.975, max_iters=10000, min_iters=50, use_sprt=True,
                   min_inlier_ratio_for_sprt=0.00001, sampler=1, neighborhood=0, neighborhood_size=4.0,
                   use_space_partitioning=True, lo_number=50, sampler_variance=0.1, solver=0, seed=None):
    """Homography by GC-RANSAC (reference: bindings.cpp:859-983, defaults :1158-1178)."""
    _check_solver(solver)
    c = _as_corr(correspondences, 4, "correspondences")
    pr, npr = _probs(probabilities)
    p = _params(0, threshold, conf, spatial_coherence_weight, max_iters, min_iters, use_sprt,
                min_inlier_ratio_for_sprt, sampler, neighborhood, neighborhood_size, use_space_partitioning,
                lo_number, sampler_variance, solver, seed)
    H = np.zeros(9)
    mask = np.zeros(c.shape[0], np.uint8)
    st = GcrStats()
    rc = _native.lib().gcr_find_homography(
        _native.dptr(c), c.shape[0], int(h1), int(w1), int(h2), int(w2), _native.dptr(pr), npr,
        ctypes.byref(p), _native.dptr(H), _native.u8ptr(mask), ctypes.byref(st))
    return _finish(rc, H, mask, st)


def findFundamentalMatrix(correspondences, h1, w1, h2, w2, probabilities, threshold=1.0, conf=0.99,
                          spatial_coherence_weight=0.0, max_iters=10000, min_iters=50, use_sprt=True,
                          min_inlier_ratio_for_sprt=0.00001, sampler=1, neighborhood=1, neighborhood_size=20.0,
                          lo_number=50, sampler_variance=0.1, solver=0, seed=None):
    """Fundamental matrix by GC-RANSAC (reference: bindings.cpp:279-400, defaults :1003-1022)."""
    _check_solver(solver)
    c = _as_corr(correspondences, 7, "x1y1")
    pr, npr = _probs(probabilities)
    p = _params(1, threshold, conf, spatial_coherence_weight, max_iters, min_iters, use_sprt,
                min_inlier_ratio_for_sprt, sampler, neighborhood, neighborhood_size, False, lo_number,
                sampler_variance, solver, seed)
    F = np.zeros(9)
    mask = np.zeros(c.shape[0], np.uint8)
    st = GcrStats()
    rc = _native.lib().gcr_find_fundamental_matrix(
        _native.dptr(c), c.shape[0], int(h1), int(w1), int(h2), int(w2), _native.dptr(pr), npr,
        ctypes.byref(p), _native.dptr(F), _native.u8ptr(mask), ctypes.byref(st))
    return _finish(rc, F, mask, st)


def findEssentialMatrix(correspondences, K1, K2, h1, w1, h2, w2, probabilities, threshold=1.0, conf=0.99,
                        spatial_coherence_weight=0.1, max_iters=10000, min_iters=50, use_sprt=True,
                        min_inlier_ratio_for_sprt=0.01, sampler=1, neighborhood=1, neighborhood_size=20.0,
                        lo_number=50, sampler_variance=0.1, solver=0, seed=None):
    """Essential matrix by GC-RANSAC (reference: bindings.cpp:710-857, defaults :1091-1112)."""
    _check_solver(solver)
    c = _as_corr(correspondences, 5, "correspondences")
    k1 = np.ascontiguousarray(np.asarray(K1, dtype=np.float64))
    k2 = np.ascontiguousarray(np.asarray(K2, dtype=np.float64))
    if k1.shape != (3, 3):
        raise ValueError("K1 shape should be [3x3]")
    if k2.shape != (3, 3):
        raise ValueError("K2 shape should be [3x3]")
    pr, npr = _probs(probabilities)
    p = _params(2, threshold, conf, spatial_coherence_weight, max_iters, min_iters, use_sprt,
                min_inlier_ratio_for_sprt, sampler, neighborhood, neighborhood_size, False, lo_number,
                sampler_variance, solver, seed)
    E = np.zeros(9)
    mask = np.zeros(c.shape[0], np.uint8)
    st = GcrStats()
    rc = _native.lib().gcr_find_essential_matrix(
        _native.dptr(c), c.shape[0], _native.dptr(k1), _native.dptr(k2), int(h1), int(w1), int(h2), int(w2),
        _native.dptr(pr), npr, ctypes.byref(p), _native.dptr(E), _native.u8ptr(mask), ctypes.byref(st))
    return _finish(rc, E, mask, st)


def findBatch(problems):
    """Solve many independent pairs in one engine run.

    ``problems`` is a list of dicts with keys ``kind`` ('homography',
    'fundamental' or 'essential'), ``correspondences``, ``h1 w1 h2 w2``,
    optional ``K1 K2`` (essential) and any keyword of the find* functions.
    Returns a list of ``(model or None, mask)``.  A pair that failed on its
    own (its min-cut gave up) returns ``(None, all False)`` and a negative
    ``status`` in ``last_batch_stats()``; the other pairs' results stand.
    """
    kinds_map = {"homography": 0, "fundamental": 1, "essential": 2}
    n = len(problems)
    kinds = np.zeros(n, np.int32)
    sizes = np.zeros(4 * n, np.int32)
    offsets = np.zeros(n + 1, np.int64)
    Ks = np.zeros(18 * n)
    params = (GcrParams * n)()
    corrs = []
    for i, pb in enumerate(problems):
        k = kinds_map[pb["kind"]]
        kinds[i] = k
        c = _as_corr(pb["correspondences"], (4, 7, 5)[k], "correspondences")
        corrs.append(c)
        offsets[i + 1] = offsets[i] + c.shape[0]
        sizes[4 * i:4 * i + 4] = [pb["h1"], pb["w1"], pb["h2"], pb["w2"]]
        if k == 2:
            Ks[18 * i:18 * i + 9] = np.asarray(pb["K1"], np.float64).ravel()
            Ks[18 * i + 9:18 * i + 18] = np.asarray(pb["K2"], np.float64).ravel()
        defaults = dict(threshold=1.0, conf=0.99, spatial_coherence_weight=(0.975, 0.0, 0.1)[k], max_iters=10000,
                        min_iters=50, use_sprt=True, min_inlier_ratio_for_sprt=(1e-5, 1e-5, 0.01)[k], sampler=1,
                        neighborhood=(0, 1, 1)[k], neighborhood_size=(4.0, 20.0, 20.0)[k],
                        use_space_partitioning=False, lo_number=50, sampler_variance=0.1, solver=0, seed=None)
        for key in defaults:
            if key in pb:
                defaults[key] = pb[key]
        _check_solver(defaults["solver"])
        params[i] = _params(k, **defaults)
    corr = np.ascontiguousarray(np.concatenate(corrs, axis=0))
    models = np.zeros(9 * n)
    masks = np.zeros(int(offsets[-1]), np.uint8)
    ninl = np.zeros(n, np.int32)
    stats = (GcrStats * n)()
    rc = _native.lib().gcr_find_batch(
        n, kinds.ctypes.data_as(ctypes.c_void_p), offsets.ctypes.data_as(ctypes.c_void_p), _native.dptr(corr), 0,
        sizes.ctypes.data_as(ctypes.c_void_p), _native.dptr(Ks), params, _native.dptr(models),
        _native.u8ptr(masks), ninl.ctypes.data_as(ctypes.c_void_p), stats)
    _native.raise_for(rc)
    global _last_batch_stats
    _last_batch_stats = [dict(stats[i].as_dict(), status=int(ninl[i])) for i in range(n)]
    out = []
    for i in range(n):
        m = masks[offsets[i]:offsets[i + 1]].astype(bool)
        out.append((None if ninl[i] <= 0 else models[9 * i:9 * i + 9].reshape(3, 3).copy(), m))
    return out
