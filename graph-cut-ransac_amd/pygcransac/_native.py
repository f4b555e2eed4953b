"""ctypes binding of libgcr_hip.so (include/gcr.h)."""
import ctypes
import os

import numpy as np

# A batch runs as up to four concurrent groups, each on its own HIP stream,
# beside the guidance stream.  HIP maps a process's streams onto
# GPU_MAX_HW_QUEUES hardware queues (4 by default), and streams that share a
# queue run one after the other, so the engine asks for 8 unless the caller
# chose a value.  HIP reads the variable when it initialises, so this holds
# when no HIP work ran in the process before this module was imported.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "libgcr_hip.so")

GCR_OK = 0
GCR_ERR_INVALID_ARGUMENT = -1
GCR_ERR_UNSUPPORTED = -2
GCR_ERR_DEVICE = -3
GCR_ERR_OUT_OF_MEMORY = -4
GCR_ERR_INTERNAL = -5

# Every symbol include/gcr.h declares.
EXPORTS = (
    "gcr_default_params",
    "gcr_find_homography",
    "gcr_find_fundamental_matrix",
    "gcr_find_essential_matrix",
    "gcr_find_batch",
    "gcr_score_models_noalloc",
    "gcr_run",
    "gcr_score_models",
    "gcr_solve_minimal",
    "gcr_sampler_draws",
    "gcr_last_batch_groups",
    "gcr_labeling",
    "gcr_fit_nonminimal",
    "gcr_profile_enable",
    "gcr_profile_reset",
    "gcr_profile_read",
    "gcr_last_error",
    "gcr_device_count",
    "gcr_mlp_num_params",
    "gcr_mlp_features",
    "gcr_mlp_predict",
    "gcr_mlp_gradient",
    "gcr_mlp_train",
    "gcr_find_batch_guided",
)


class GcrParams(ctypes.Structure):
    _fields_ = [
        ("threshold", ctypes.c_double),
        ("confidence", ctypes.c_double),
        ("spatial_coherence_weight", ctypes.c_double),
        ("max_iters", ctypes.c_int64),
        ("min_iters", ctypes.c_int64),
        ("use_sprt", ctypes.c_int32),
        ("min_inlier_ratio_for_sprt", ctypes.c_double),
        ("sampler", ctypes.c_int32),
        ("neighborhood", ctypes.c_int32),
        ("neighborhood_size", ctypes.c_double),
        ("use_space_partitioning", ctypes.c_int32),
        ("lo_number", ctypes.c_int32),
        ("sampler_variance", ctypes.c_double),
        ("solver", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("time_limit", ctypes.c_double),
        ("sample_fn", ctypes.c_void_p),
        ("sample_ctx", ctypes.c_void_p),
    ]


class GcrStats(ctypes.Structure):
    _fields_ = [
        ("iteration_number", ctypes.c_int64),
        ("local_optimization_number", ctypes.c_int64),
        ("graph_cut_number", ctypes.c_int64),
        ("num_inliers", ctypes.c_int64),
        ("score", ctypes.c_double),
        ("processing_time", ctypes.c_double),
        ("candidates_solved", ctypes.c_int64),
        ("models_scored", ctypes.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def library_path():
    """Path of the in-tree engine library."""
    return LIB_PATH


def lib():
    """Load the engine library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libgcr_hip.so is not built (run __graft_entry__.build() or make in %s)" % _PKG_ROOT)
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, d = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        L.gcr_default_params.argtypes = [vp, i32]
        L.gcr_default_params.restype = None
        L.gcr_find_homography.argtypes = [vp, i64, i32, i32, i32, i32, vp, i64, vp, vp, vp, vp]
        L.gcr_find_homography.restype = i32
        L.gcr_find_fundamental_matrix.argtypes = [vp, i64, i32, i32, i32, i32, vp, i64, vp, vp, vp, vp]
        L.gcr_find_fundamental_matrix.restype = i32
        L.gcr_find_essential_matrix.argtypes = [vp, i64, vp, vp, i32, i32, i32, i32, vp, i64, vp, vp, vp, vp]
        L.gcr_find_essential_matrix.restype = i32
        L.gcr_find_batch.argtypes = [i32, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp]
        L.gcr_find_batch.restype = i32
        L.gcr_run.argtypes = [i32, vp, i64, vp, vp, i32, i32, vp, vp, vp, vp]
        L.gcr_run.restype = i32
        L.gcr_score_models.argtypes = [i32, vp, i64, vp, i64, d, vp, vp, vp]
        L.gcr_score_models.restype = i32
        L.gcr_score_models_noalloc.argtypes = [i32, vp, i64, vp, i64, d, vp, vp, vp, vp]
        L.gcr_score_models_noalloc.restype = i32
        L.gcr_solve_minimal.argtypes = [i32, vp, i64, vp, i64, vp, vp]
        L.gcr_solve_minimal.restype = i32
        L.gcr_sampler_draws.argtypes = [i32, ctypes.c_uint32, i64, i32, vp, i32, vp]
        L.gcr_sampler_draws.restype = i32
        L.gcr_last_batch_groups.argtypes = []
        L.gcr_last_batch_groups.restype = i32
        L.gcr_labeling.argtypes = [i32, vp, i64, vp, d, d, vp, i64, vp]
        L.gcr_labeling.restype = i32
        L.gcr_fit_nonminimal.argtypes = [i32, vp, i64, vp, i64, vp, i32, i32, d, vp]
        L.gcr_fit_nonminimal.restype = i32
        L.gcr_profile_enable.argtypes = [i32]
        L.gcr_profile_enable.restype = None
        L.gcr_profile_reset.argtypes = []
        L.gcr_profile_reset.restype = None
        L.gcr_profile_read.argtypes = [ctypes.c_char_p, vp, vp, vp]
        L.gcr_profile_read.restype = i32
        L.gcr_last_error.argtypes = []
        L.gcr_last_error.restype = ctypes.c_char_p
        L.gcr_device_count.argtypes = []
        L.gcr_device_count.restype = i32
        f32, u64 = ctypes.c_float, ctypes.c_uint64
        L.gcr_mlp_num_params.argtypes = []
        L.gcr_mlp_num_params.restype = i32
        L.gcr_mlp_features.argtypes = [i32, vp, vp, vp, vp, vp]
        L.gcr_mlp_features.restype = i32
        L.gcr_mlp_predict.argtypes = [vp, vp, i64, vp, vp]
        L.gcr_mlp_predict.restype = i32
        L.gcr_mlp_gradient.argtypes = [vp, vp, vp, i64, vp, vp]
        L.gcr_mlp_gradient.restype = i32
        L.gcr_mlp_train.argtypes = [vp, vp, vp, i64, vp, vp, i64, i32, i32, f32, u64, vp]
        L.gcr_mlp_train.restype = i32
        L.gcr_find_batch_guided.argtypes = [i32, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.gcr_find_batch_guided.restype = i32
        _lib = L
    return _lib


def dptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def u8ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def raise_for(rc):
    if rc >= 0:
        return
    msg = lib().gcr_last_error().decode("utf-8", "replace")
    if rc == GCR_ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    if rc == GCR_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == GCR_ERR_OUT_OF_MEMORY:
        raise MemoryError(msg)
    raise RuntimeError("gcr error %d: %s" % (rc, msg))


def score_models(kind, pts, models, sq_thr, want_masks=False):
    """Device MSAC scores (counts, values[, masks]) of k models."""
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    models = np.ascontiguousarray(models, dtype=np.float64).reshape(-1, 9)
    k = models.shape[0]
    counts = np.zeros(k, np.int64)
    values = np.zeros(k)
    masks = np.zeros((k, pts.shape[0]), np.uint8) if want_masks else None
    rc = lib().gcr_score_models(int(kind), dptr(pts), pts.shape[0], dptr(models), k, float(sq_thr), dptr(counts),
                                dptr(values), dptr(masks))
    raise_for(rc)
    return (counts, values, masks) if want_masks else (counts, values)


def solve_minimal(kind, pts, samples):
    """kind 0: (models c x 9, status c); kind 1: (models c x 3 x 9, status c x 3);
    kind 2: (models c x 10 x 9, status c x 10)."""
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    samples = np.ascontiguousarray(samples, dtype=np.int32)
    c = samples.shape[0]
    slots = (1, 3, 10)[kind]
    models = np.zeros((c, 9)) if kind == 0 else np.zeros((c, slots, 9))
    status = np.zeros(c, np.int32) if kind == 0 else np.zeros((c, slots), np.int32)
    rc = lib().gcr_solve_minimal(int(kind), dptr(pts), pts.shape[0], dptr(samples), c, dptr(models), dptr(status))
    raise_for(rc)
    return models, status


def sampler_draws(sampler, seed, n, m, chunks):
    """Consecutive chunks of the uniform (0) or PROSAC (1) sampler stream
    over n points, drawn on the device: (sum(chunks), m) int32 indices."""
    ch = np.ascontiguousarray(chunks, dtype=np.int64)
    out = np.zeros((int(ch.sum()), m), np.int32)
    rc = lib().gcr_sampler_draws(int(sampler), int(seed) & 0xFFFFFFFF, int(n), int(m), dptr(ch), len(ch), dptr(out))
    raise_for(rc)
    return out


def labeling(kind, pts, model, lam, threshold, cell_sizes, divisions):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    model = np.ascontiguousarray(model, dtype=np.float64).ravel()
    cell = np.ascontiguousarray(cell_sizes, dtype=np.float64)
    sink = np.zeros(pts.shape[0], np.uint8)
    rc = lib().gcr_labeling(int(kind), dptr(pts), pts.shape[0], dptr(model), float(lam), float(threshold), dptr(cell),
                            int(divisions), dptr(sink))
    raise_for(rc)
    return sink


def fit_nonminimal(kind, pts, sample, weights=None, weight_mode=0, normalize=1, threshold=1.0):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    sample = np.ascontiguousarray(sample, dtype=np.int32)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64).ravel()
    out = np.zeros((40, 9))
    k = lib().gcr_fit_nonminimal(int(kind), dptr(pts), pts.shape[0], dptr(sample), sample.size, dptr(w),
                                 int(weight_mode), int(normalize), float(threshold), dptr(out))
    raise_for(k)
    return out[:k].copy()


def profile_enable(on=True):
    lib().gcr_profile_enable(1 if on else 0)


def profile_reset():
    lib().gcr_profile_reset()


def profile_read(family):
    """(total_ms, launches, work) of a kernel family, or None."""
    ms = ctypes.c_double()
    n = ctypes.c_int64()
    w = ctypes.c_double()
    rc = lib().gcr_profile_read(family.encode(), ctypes.byref(ms), ctypes.byref(n), ctypes.byref(w))
    if rc != 0:
        return None
    return ms.value, n.value, w.value
