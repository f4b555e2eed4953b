"""Learned guidance: a per-correspondence inlier-probability network whose
ordering drives the PROSAC sampler.

The reference's PROSAC mode (src/pygcransac/include/samplers/
prosac_sampler.h:51-237) expects the correspondences sorted by a quality
score; the drivers pass the caller's order through
(src/pygcransac/src/gcransac_python.cpp:4344-4359).  This module computes
that score on the device with a small network (8-128-128-128-1, GELU, BF16
weights, FP32 accumulation on the matrix cores), trains it on the device
from GC-RANSAC inlier masks, and runs guided batches (features -> logits ->
per-pair stable descending sort -> batched GC-RANSAC with PROSAC on the
sorted rows -> masks in the caller's order).  Full-set MSAC scoring still
judges every model.

Everything runs in libgcr_hip.so through its C ABI (include/gcr.h,
"learned guidance"); there is no host fallback.
"""
import ctypes
import os

import numpy as np

from . import _native
from ._native import GcrParams, GcrStats

NUM_PARAMS = 34305
NUM_FEATURES = 8
OFF = dict(W1=0, b1=1024, W2=1152, b2=17536, W3=17664, b3=34048, w4=34176, b4=34304)
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "guided_mlp.npz")

_KINDS = {"homography": 0, "fundamental": 1, "essential": 2}


def _vp(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def init_params(seed=0):
    """He-normal weights (fan-in), zero biases; the padded input column of W1
    is zero and stays zero under training (its gradient is zero)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NUM_PARAMS, np.float32)
    w1 = rng.normal(size=(128, 8)) * np.sqrt(2.0 / 7.0)
    w1[:, 7] = 0.0
    p[OFF["W1"]:OFF["b1"]] = w1.ravel()
    p[OFF["W2"]:OFF["b2"]] = (rng.normal(size=(128, 128)) * np.sqrt(2.0 / 128)).ravel()
    p[OFF["W3"]:OFF["b3"]] = (rng.normal(size=(128, 128)) * np.sqrt(2.0 / 128)).ravel()
    p[OFF["w4"]:OFF["b4"]] = rng.normal(size=128) * np.sqrt(1.0 / 128)
    return p


def save(path, params):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez(path, params=np.asarray(params, np.float32))


def load(path=WEIGHTS_PATH):
    with np.load(path, allow_pickle=False) as z:
        p = np.ascontiguousarray(z["params"], dtype=np.float32)
    if p.shape != (NUM_PARAMS,):
        raise ValueError("network parameter file has %s entries, expected %d" % (p.shape, NUM_PARAMS))
    return p


def _pack_pairs(pairs):
    n = len(pairs)
    offsets = np.zeros(n + 1, np.int64)
    sizes = np.zeros(4 * n, np.int32)
    corrs, ratios = [], []
    for i, pb in enumerate(pairs):
        c = np.ascontiguousarray(np.asarray(pb["correspondences"], np.float64))
        if c.ndim != 2 or c.shape[1] != 4 or c.shape[0] < 1:
            raise ValueError("correspondences should be an array with dims [n,4], n>=1")
        r = np.ascontiguousarray(np.asarray(pb["ratio"], np.float64).ravel())
        if r.shape[0] != c.shape[0]:
            raise ValueError("ratio needs one value per correspondence")
        corrs.append(c)
        ratios.append(r)
        offsets[i + 1] = offsets[i] + c.shape[0]
        sizes[4 * i:4 * i + 4] = [pb["h1"], pb["w1"], pb["h2"], pb["w2"]]
    return offsets, sizes, np.concatenate(corrs, 0), np.concatenate(ratios)


def features(pairs):
    """Features of each pair (list of dicts: correspondences, ratio, h1 w1
    h2 w2) -> list of (n, 8) float32 arrays."""
    offsets, sizes, corr, ratio = _pack_pairs(pairs)
    out = np.zeros((int(offsets[-1]), NUM_FEATURES), np.float32)
    rc = _native.lib().gcr_mlp_features(len(pairs), _vp(offsets), _vp(corr), _vp(sizes), _vp(ratio), _vp(out))
    _native.raise_for(rc)
    return [out[offsets[i]:offsets[i + 1]] for i in range(len(pairs))]


def predict(params, feat):
    """(logits, probabilities) of feature rows."""
    params = np.ascontiguousarray(params, np.float32)
    feat = np.ascontiguousarray(feat, np.float32).reshape(-1, NUM_FEATURES)
    logit = np.zeros(feat.shape[0], np.float32)
    prob = np.zeros(feat.shape[0], np.float32)
    rc = _native.lib().gcr_mlp_predict(_vp(params), _vp(feat), feat.shape[0], _vp(logit), _vp(prob))
    _native.raise_for(rc)
    return logit, prob


def gradient(params, feat, labels):
    """(mean BCE loss, gradient) from one device forward + backward pass."""
    params = np.ascontiguousarray(params, np.float32)
    feat = np.ascontiguousarray(feat, np.float32).reshape(-1, NUM_FEATURES)
    labels = np.ascontiguousarray(labels, np.float32).ravel()
    grad = np.zeros(NUM_PARAMS, np.float32)
    loss = ctypes.c_float(0.0)
    rc = _native.lib().gcr_mlp_gradient(_vp(params), _vp(feat), _vp(labels), feat.shape[0], _vp(grad),
                                        ctypes.byref(loss))
    _native.raise_for(rc)
    return float(loss.value), grad


class Trainer:
    """Adam state of one network; train() runs whole steps on the device."""

    def __init__(self, params=None, seed=0):
        self.params = np.ascontiguousarray(init_params(seed) if params is None else params, np.float32).copy()
        self.m = np.zeros(NUM_PARAMS, np.float32)
        self.v = np.zeros(NUM_PARAMS, np.float32)
        self.step = 0

    def train(self, feat, labels, steps, batch=8192, lr=1e-3, seed=0):
        feat = np.ascontiguousarray(feat, np.float32).reshape(-1, NUM_FEATURES)
        labels = np.ascontiguousarray(labels, np.float32).ravel()
        losses = np.zeros(max(int(steps), 1), np.float32)
        rc = _native.lib().gcr_mlp_train(_vp(self.params), _vp(self.m), _vp(self.v), self.step, _vp(feat),
                                         _vp(labels), feat.shape[0], int(batch), int(steps), float(lr),
                                         int(seed) & 0xFFFFFFFFFFFFFFFF, _vp(losses))
        _native.raise_for(rc)
        self.step += int(steps)
        return losses[:int(steps)]


def _batch_params(problems):
    from . import _params, _check_solver
    n = len(problems)
    kinds = np.zeros(n, np.int32)
    Ks = np.zeros(18 * n)
    params = (GcrParams * n)()
    for i, pb in enumerate(problems):
        k = _KINDS[pb["kind"]]
        kinds[i] = k
        if k == 2:
            Ks[18 * i:18 * i + 9] = np.asarray(pb["K1"], np.float64).ravel()
            Ks[18 * i + 9:18 * i + 18] = np.asarray(pb["K2"], np.float64).ravel()
        d = dict(threshold=1.0, conf=0.99, spatial_coherence_weight=(0.975, 0.0, 0.1)[k], max_iters=10000,
                 min_iters=50, use_sprt=True, min_inlier_ratio_for_sprt=(1e-5, 1e-5, 0.01)[k], sampler=1,
                 neighborhood=(0, 1, 1)[k], neighborhood_size=(4.0, 20.0, 20.0)[k], use_space_partitioning=False,
                 lo_number=50, sampler_variance=0.1, solver=0, seed=None)
        for key in d:
            if key in pb:
                d[key] = pb[key]
        _check_solver(d["solver"])
        params[i] = _params(k, **d)
    return kinds, Ks, params


def find_batch_guided(problems, params, return_logits=False):
    """Guided batch: ``problems`` as for pygcransac.findBatch plus a
    ``ratio`` array per pair (the descriptor-ratio score); ``params`` the
    network.  Returns a list of ``(model or None, mask)`` in the caller's row
    order (and the logits when asked).  The sampler is PROSAC over the
    network's order whatever ``sampler`` says."""
    offsets, sizes, corr, ratio = _pack_pairs(problems)
    kinds, Ks, gp = _batch_params(problems)
    n = len(problems)
    net = np.ascontiguousarray(params, np.float32)
    models = np.zeros(9 * n)
    masks = np.zeros(int(offsets[-1]), np.uint8)
    ninl = np.zeros(n, np.int32)
    stats = (GcrStats * n)()
    logits = np.zeros(int(offsets[-1]), np.float32) if return_logits else None
    rc = _native.lib().gcr_find_batch_guided(n, _vp(kinds), _vp(offsets), _vp(corr), 0, _vp(sizes), _vp(Ks),
                                             _vp(ratio), _vp(net), gp, _vp(models), _vp(masks), _vp(ninl), stats,
                                             _vp(logits))
    _native.raise_for(rc)
    out = []
    for i in range(n):
        m = masks[offsets[i]:offsets[i + 1]].astype(bool)
        out.append((None if ninl[i] <= 0 else models[9 * i:9 * i + 9].reshape(3, 3).copy(), m))
    global last_stats
    last_stats = [stats[i].as_dict() for i in range(n)]
    if return_logits:
        return out, [logits[offsets[i]:offsets[i + 1]] for i in range(n)]
    return out


last_stats = None
