"""ctypes wrapper of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Used by tests/ as the parity checker and by bench.py's cpu_baseline leg.
The engine never imports this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OracleParams(ctypes.Structure):
    _fields_ = [
        ("threshold", ctypes.c_double),
        ("confidence", ctypes.c_double),
        ("spatial_coherence_weight", ctypes.c_double),
        ("max_iters", ctypes.c_int64),
        ("min_iters", ctypes.c_int64),
        ("sampler", ctypes.c_int32),
        ("lo_number", ctypes.c_int32),
        ("neighborhood_size", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("probabilities", ctypes.c_void_p),
        ("sampler_variance", ctypes.c_double),
        ("neighborhood", ctypes.c_int32),
        ("use_sprt", ctypes.c_int32),
        ("min_inlier_ratio_for_sprt", ctypes.c_double),
        ("use_space_partitioning", ctypes.c_int32),
    ]


class OracleStats(ctypes.Structure):
    _fields_ = [
        ("iteration_number", ctypes.c_int64),
        ("local_optimization_number", ctypes.c_int64),
        ("graph_cut_number", ctypes.c_int64),
        ("num_inliers", ctypes.c_int64),
        ("score", ctypes.c_double),
        ("seconds", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def params(threshold, conf=0.99, spatial=0.975, max_iters=10000, min_iters=50, sampler=0, lo_number=50,
           neighborhood_size=8.0, seed=0, probabilities=None, sampler_variance=0.1, neighborhood=0, use_sprt=False,
           min_inlier_ratio_for_sprt=0.00001, use_space_partitioning=False):
    """probabilities: per point (samplers 3 and 4); the array is kept alive
    on the returned structure.  neighborhood 0: grid of neighborhood_size
    divisions per axis; 1: radius graph of radius neighborhood_size."""
    pr = None if probabilities is None else np.ascontiguousarray(probabilities, np.float64)
    p = OracleParams(threshold, conf, spatial, max_iters, min_iters, sampler, lo_number, neighborhood_size, seed,
                     None if pr is None else pr.ctypes.data_as(ctypes.c_void_p), sampler_variance, neighborhood,
                     int(bool(use_sprt)), min_inlier_ratio_for_sprt, int(bool(use_space_partitioning)))
    p._probabilities = pr
    return p


def find_homography(corr, h1, w1, h2, w2, p):
    corr = np.ascontiguousarray(corr, np.float64)
    n = corr.shape[0]
    H = np.zeros(9)
    mask = np.zeros(n, np.uint8)
    st = OracleStats()
    L = lib()
    L.oracle_find_homography.restype = ctypes.c_int
    rc = L.oracle_find_homography(_p(corr), ctypes.c_int64(n), h1, w1, h2, w2, ctypes.byref(p), _p(H), _p(mask),
                                  ctypes.byref(st))
    return rc, H.reshape(3, 3), mask.astype(bool), st.as_dict()


def find_fundamental(corr, h1, w1, h2, w2, p):
    corr = np.ascontiguousarray(corr, np.float64)
    n = corr.shape[0]
    F = np.zeros(9)
    mask = np.zeros(n, np.uint8)
    st = OracleStats()
    L = lib()
    L.oracle_find_fundamental_matrix.restype = ctypes.c_int
    rc = L.oracle_find_fundamental_matrix(_p(corr), ctypes.c_int64(n), h1, w1, h2, w2, ctypes.byref(p), _p(F),
                                          _p(mask), ctypes.byref(st))
    return rc, F.reshape(3, 3), mask.astype(bool), st.as_dict()


def find_essential(corr, K1, K2, h1, w1, h2, w2, p):
    corr = np.ascontiguousarray(corr, np.float64)
    K1 = np.ascontiguousarray(K1, np.float64)
    K2 = np.ascontiguousarray(K2, np.float64)
    n = corr.shape[0]
    E = np.zeros(9)
    mask = np.zeros(n, np.uint8)
    st = OracleStats()
    L = lib()
    L.oracle_find_essential_matrix.restype = ctypes.c_int
    rc = L.oracle_find_essential_matrix(_p(corr), ctypes.c_int64(n), _p(K1), _p(K2), h1, w1, h2, w2,
                                        ctypes.byref(p), _p(E), _p(mask), ctypes.byref(st))
    return rc, E.reshape(3, 3), mask.astype(bool), st.as_dict()


def score(kind, pts, models, sq_thr):
    pts = np.ascontiguousarray(pts, np.float64)
    models = np.ascontiguousarray(models, np.float64).reshape(-1, 9)
    k = models.shape[0]
    counts = np.zeros(k, np.int64)
    values = np.zeros(k)
    masks = np.zeros((k, pts.shape[0]), np.uint8)
    lib().oracle_score(kind, _p(pts), ctypes.c_int64(pts.shape[0]), _p(models), ctypes.c_int64(k),
                       ctypes.c_double(sq_thr), _p(counts), _p(values), _p(masks))
    return counts, values, masks


def h4_solve(pts, samples):
    pts = np.ascontiguousarray(pts, np.float64)
    samples = np.ascontiguousarray(samples, np.int32)
    c = samples.shape[0]
    models = np.zeros((c, 9))
    status = np.zeros(c, np.int32)
    lib().oracle_h4_solve(_p(pts), ctypes.c_int64(pts.shape[0]), _p(samples), ctypes.c_int64(c), _p(models),
                          _p(status))
    return models, status


def h_nonminimal(pts, sample, weights=None):
    pts = np.ascontiguousarray(pts, np.float64)
    sample = np.ascontiguousarray(sample, np.int64)
    out = np.zeros(9)
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    L = lib()
    L.oracle_h_nonminimal.restype = ctypes.c_int
    ok = L.oracle_h_nonminimal(_p(pts), ctypes.c_int64(pts.shape[0]), _p(sample), ctypes.c_int64(sample.size), _p(w),
                               _p(out))
    return ok, out.reshape(3, 3)


def uniform_draws(seed, pool_n, m, count):
    out = np.zeros((count, m), np.int64)
    lib().oracle_uniform_draws(ctypes.c_uint32(seed), ctypes.c_int64(pool_n), ctypes.c_int32(m), ctypes.c_int64(count),
                               _p(out))
    return out


def prosac_draws(seed, n, m, count):
    out = np.zeros((count, m), np.int64)
    lib().oracle_prosac_draws(ctypes.c_uint32(seed), ctypes.c_int64(n), ctypes.c_int32(m), ctypes.c_int64(count),
                              _p(out))
    return out


def prosac_growth(n, m):
    out = np.zeros(n, np.int64)
    lib().oracle_prosac_growth(ctypes.c_int64(n), ctypes.c_int32(m), _p(out))
    return out


def grid_keys(pts, cell_sizes, divisions):
    pts = np.ascontiguousarray(pts, np.float64)
    cell = np.ascontiguousarray(cell_sizes, np.float64)
    keys = np.zeros(pts.shape[0], np.uint64)
    ne = ctypes.c_int64(0)
    lib().oracle_grid_keys(_p(pts), ctypes.c_int64(pts.shape[0]), _p(cell), ctypes.c_int64(divisions), _p(keys),
                           ctypes.byref(ne))
    return keys, ne.value


def labeling(kind, pts, model, lam, threshold, cell_sizes, divisions, want_graph=False):
    """SINK mask of GCRANSAC::labeling; with want_graph also the energy graph."""
    pts = np.ascontiguousarray(pts, np.float64)
    model = np.ascontiguousarray(model, np.float64).ravel()
    cell = np.ascontiguousarray(cell_sizes, np.float64)
    n = pts.shape[0]
    sink = np.zeros(n, np.uint8)
    tr = np.zeros(n)
    L = lib()
    L.oracle_labeling.restype = ctypes.c_int64
    ne = L.oracle_labeling(kind, _p(pts), ctypes.c_int64(n), _p(model), ctypes.c_double(lam),
                           ctypes.c_double(threshold), _p(cell), ctypes.c_int64(divisions), _p(sink), _p(tr),
                           ctypes.c_int64(0), None, None, None, None)
    if not want_graph:
        return sink.astype(bool)
    ep = np.zeros(ne, np.int32)
    eq = np.zeros(ne, np.int32)
    cpq = np.zeros(ne)
    cqp = np.zeros(ne)
    L.oracle_labeling(kind, _p(pts), ctypes.c_int64(n), _p(model), ctypes.c_double(lam), ctypes.c_double(threshold),
                      _p(cell), ctypes.c_int64(divisions), _p(sink), _p(tr), ctypes.c_int64(ne), _p(ep), _p(eq),
                      _p(cpq), _p(cqp))
    return sink.astype(bool), dict(tr_cap=tr, ep=ep, eq=eq, cap_pq=cpq, cap_qp=cqp)


def labeling_radius(kind, pts, model, lam, threshold, radius, want_graph=False):
    """SINK mask of GCRANSAC::labeling on the radius neighbourhood."""
    pts = np.ascontiguousarray(pts, np.float64)
    model = np.ascontiguousarray(model, np.float64).ravel()
    n = pts.shape[0]
    sink = np.zeros(n, np.uint8)
    tr = np.zeros(n)
    L = lib()
    L.oracle_labeling_radius.restype = ctypes.c_int64
    ne = L.oracle_labeling_radius(kind, _p(pts), ctypes.c_int64(n), _p(model), ctypes.c_double(lam),
                                  ctypes.c_double(threshold), ctypes.c_double(radius), _p(sink), _p(tr),
                                  ctypes.c_int64(0), None, None, None, None)
    if not want_graph:
        return sink.astype(bool)
    ep = np.zeros(ne, np.int32)
    eq = np.zeros(ne, np.int32)
    cpq = np.zeros(ne)
    cqp = np.zeros(ne)
    L.oracle_labeling_radius(kind, _p(pts), ctypes.c_int64(n), _p(model), ctypes.c_double(lam),
                             ctypes.c_double(threshold), ctypes.c_double(radius), _p(sink), _p(tr), ctypes.c_int64(ne),
                             _p(ep), _p(eq), _p(cpq), _p(cqp))
    return sink.astype(bool), dict(tr_cap=tr, ep=ep, eq=eq, cap_pq=cpq, cap_qp=cqp)


def radius_neighbors(pts, radius):
    """The radius graph's neighbour lists as (off, nbr): list i is
    nbr[off[i]:off[i+1]], in (distance, index) order."""
    pts = np.ascontiguousarray(pts, np.float64)
    n = pts.shape[0]
    L = lib()
    L.oracle_radius_neighbors.restype = ctypes.c_int64
    off = np.zeros(n + 1, np.int64)
    total = L.oracle_radius_neighbors(_p(pts), ctypes.c_int64(n), ctypes.c_double(radius), None, None,
                                      ctypes.c_int64(0))
    nbr = np.zeros(max(total, 1), np.int32)
    L.oracle_radius_neighbors(_p(pts), ctypes.c_int64(n), ctypes.c_double(radius), _p(off), _p(nbr),
                              ctypes.c_int64(total))
    return off, nbr[:total]


def radius_sq_distance(a, b):
    """The radius graph's float squared distance of two float32 4-vectors."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    L = lib()
    L.oracle_radius_sq_distance.restype = ctypes.c_float
    return float(L.oracle_radius_sq_distance(_p(a), _p(b)))


def mincut(tr_cap, ep, eq, cpq, cqp):
    tr_cap = np.ascontiguousarray(tr_cap, np.float64)
    ep = np.ascontiguousarray(ep, np.int32)
    eq = np.ascontiguousarray(eq, np.int32)
    cpq = np.ascontiguousarray(cpq, np.float64)
    cqp = np.ascontiguousarray(cqp, np.float64)
    n = tr_cap.size
    sink = np.zeros(n, np.uint8)
    lib().oracle_mincut(ctypes.c_int64(n), _p(tr_cap), ctypes.c_int64(ep.size), _p(ep), _p(eq), _p(cpq), _p(cqp),
                        _p(sink))
    return sink.astype(bool)


def energy_terms(kind, pts, model, lam, threshold, cell_sizes, divisions):
    pts = np.ascontiguousarray(pts, np.float64)
    model = np.ascontiguousarray(model, np.float64).ravel()
    cell = np.ascontiguousarray(cell_sizes, np.float64)
    n = pts.shape[0]
    ua = np.zeros(n)
    ub = np.zeros(n)
    L = lib()
    L.oracle_energy_terms.restype = ctypes.c_int64
    args = (kind, _p(pts), ctypes.c_int64(n), _p(model), ctypes.c_double(lam), ctypes.c_double(threshold), _p(cell),
            ctypes.c_int64(divisions), _p(ua), _p(ub))
    ne = L.oracle_energy_terms(*args, ctypes.c_int64(0), None, None, None, None, None, None)
    ep = np.zeros(ne, np.int32)
    eq = np.zeros(ne, np.int32)
    ta, tb, tc, td = (np.zeros(ne) for _ in range(4))
    L.oracle_energy_terms(*args, ctypes.c_int64(ne), _p(ep), _p(eq), _p(ta), _p(tb), _p(tc), _p(td))
    return dict(ua=ua, ub=ub, ep=ep, eq=eq, ta=ta, tb=tb, tc=tc, td=td)


def energy_mincut(t):
    n = t["ua"].size
    ne = t["ep"].size
    sink = np.zeros(n, np.uint8)
    c = {k: np.ascontiguousarray(t[k], np.int32 if k in ("ep", "eq") else np.float64) for k in t}
    lib().oracle_energy_mincut(ctypes.c_int64(n), _p(c["ua"]), _p(c["ub"]), ctypes.c_int64(ne), _p(c["ep"]),
                               _p(c["eq"]), _p(c["ta"]), _p(c["tb"]), _p(c["tc"]), _p(c["td"]), _p(sink))
    return sink.astype(bool)


def f7_solve(pts, samples):
    pts = np.ascontiguousarray(pts, np.float64)
    samples = np.ascontiguousarray(samples, np.int32)
    c = samples.shape[0]
    models = np.zeros((c, 3, 9))
    status = np.zeros((c, 3), np.int32)
    lib().oracle_f7_solve(_p(pts), ctypes.c_int64(pts.shape[0]), _p(samples), ctypes.c_int64(c), _p(models),
                          _p(status))
    return models, status


def fit_nonminimal(kind, pts, sample, weights=None, weight_mode=0, normalize=1, threshold=1.0):
    pts = np.ascontiguousarray(pts, np.float64)
    sample = np.ascontiguousarray(sample, np.int64)
    out = np.zeros((10, 9))
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    L = lib()
    L.oracle_fit_nonminimal.restype = ctypes.c_int
    k = L.oracle_fit_nonminimal(kind, _p(pts), ctypes.c_int64(pts.shape[0]), _p(sample), ctypes.c_int64(sample.size),
                                _p(w), weight_mode, normalize, ctypes.c_double(threshold), _p(out))
    return out[:k].copy()


# ---- essential-matrix components ------------------------------------------------

def sturm10(coeffs):
    c = np.ascontiguousarray(coeffs, np.float64)
    roots = np.zeros(10)
    L = lib()
    L.oracle_sturm10.restype = ctypes.c_int
    k = L.oracle_sturm10(_p(c), _p(roots))
    return roots[:k].copy()


def e5_solve(pts, samples, orient=True):
    pts = np.ascontiguousarray(pts, np.float64)
    samples = np.ascontiguousarray(samples, np.int32)
    c = samples.shape[0]
    models = np.zeros((c, 10, 9))
    nm = np.zeros(c, np.int32)
    lib().oracle_e5_solve(_p(pts), ctypes.c_int64(pts.shape[0]), _p(samples), ctypes.c_int64(c),
                          ctypes.c_int32(1 if orient else 0), _p(models), _p(nm))
    return models, nm


def e_nonminimal(pts, sample, weights=None):
    pts = np.ascontiguousarray(pts, np.float64)
    sample = np.ascontiguousarray(sample, np.int64)
    out = np.zeros((10, 9))
    L = lib()
    L.oracle_e_nonminimal.restype = ctypes.c_int
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    k = L.oracle_e_nonminimal(_p(pts), ctypes.c_int64(pts.shape[0]), _p(sample), ctypes.c_int64(sample.size),
                              None if w is None else _p(w), _p(out))
    return out[:k].copy()


def e_bundle(pts, sample, w_pos=None):
    pts = np.ascontiguousarray(pts, np.float64)
    sample = np.ascontiguousarray(sample, np.int64)
    init = np.zeros((10, 9))
    out = np.zeros((40, 9))
    ni = ctypes.c_int32(0)
    L = lib()
    L.oracle_e_bundle.restype = ctypes.c_int
    w = None if w_pos is None else np.ascontiguousarray(w_pos, np.float64)
    k = L.oracle_e_bundle(_p(pts), ctypes.c_int64(pts.shape[0]), _p(sample), ctypes.c_int64(sample.size),
                          None if w is None else _p(w), _p(init), ctypes.byref(ni), _p(out))
    return init[:ni.value].copy(), out[:min(k, 40)].copy()


def motion_from_essential(E, x1, x2):
    E = np.ascontiguousarray(E, np.float64).ravel()
    x1 = np.ascontiguousarray(x1, np.float64)
    x2 = np.ascontiguousarray(x2, np.float64)
    out = np.zeros((4, 7))
    L = lib()
    L.oracle_motion_from_essential.restype = ctypes.c_int
    k = L.oracle_motion_from_essential(_p(E), _p(x1), _p(x2), _p(out))
    return out[:k].copy()


def normalise_correspondences(corr, K1, K2):
    corr = np.ascontiguousarray(corr, np.float64)
    K1 = np.ascontiguousarray(K1, np.float64)
    K2 = np.ascontiguousarray(K2, np.float64)
    out = np.zeros_like(corr)
    lib().oracle_normalise_correspondences(_p(corr), ctypes.c_int64(corr.shape[0]), _p(K1), _p(K2), _p(out))
    return out


def e_post_bundle(pts, sample, model, trunc_thr):
    pts = np.ascontiguousarray(pts, np.float64)
    sample = np.ascontiguousarray(sample, np.int64)
    model = np.ascontiguousarray(model, np.float64).ravel()
    out = np.zeros((40, 9))
    L = lib()
    L.oracle_e_post_bundle.restype = ctypes.c_int
    k = L.oracle_e_post_bundle(_p(pts), ctypes.c_int64(pts.shape[0]), _p(sample), ctypes.c_int64(sample.size),
                               _p(model), ctypes.c_double(trunc_thr), _p(out))
    return out[:min(k, 40)].copy()


def run(kind, pts, grid_pts, cell_sizes, p, flags=0):
    """GCRANSAC::run as the template API calls it (no driver stage)."""
    pts = np.ascontiguousarray(pts, np.float64)
    gp = None if grid_pts is None else np.ascontiguousarray(grid_pts, np.float64)
    cs = np.ascontiguousarray(cell_sizes, np.float64)
    n = pts.shape[0]
    M = np.zeros(9)
    mask = np.zeros(n, np.uint8)
    st = OracleStats()
    L = lib()
    L.oracle_run.restype = ctypes.c_int
    rc = L.oracle_run(ctypes.c_int32(kind), _p(pts), ctypes.c_int64(n), _p(gp), _p(cs), ctypes.c_int32(flags),
                      ctypes.byref(p), _p(M), _p(mask), ctypes.byref(st))
    return rc, M.reshape(3, 3), mask.astype(bool), st.as_dict()


# ---- learned inlier-probability network (oracle_mlp.hpp) -----------------

def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def mlp_num_params():
    return int(lib().oracle_mlp_num_params())


def mlp_features(corr, h1, w1, h2, w2, ratio):
    corr = np.ascontiguousarray(corr, dtype=np.float64)
    ratio = np.ascontiguousarray(ratio, dtype=np.float64)
    n = len(corr)
    out = np.zeros((n, 8), np.float32)
    L = lib()
    L.oracle_mlp_features.restype = None
    L.oracle_mlp_features(_p(corr), ctypes.c_int64(n), h1, w1, h2, w2, _p(ratio), _p(out))
    return out


def mlp_knn(corr):
    corr = np.ascontiguousarray(corr, dtype=np.float64)
    out = np.zeros((len(corr), 8), np.int32)
    L = lib()
    L.oracle_mlp_knn.restype = None
    L.oracle_mlp_knn(_p(corr), ctypes.c_int64(len(corr)), _p(out))
    return out


def mlp_logits(params, feat, exact=False):
    params, feat = _f32(params), _f32(feat)
    out = np.zeros(len(feat), np.float64)
    L = lib()
    L.oracle_mlp_logits.restype = None
    L.oracle_mlp_logits(_p(params), _p(feat), ctypes.c_int64(len(feat)), int(exact), _p(out))
    return out


def mlp_loss_grad(params, feat, labels, exact=False, params_d=None, want_grad=True):
    params, feat, labels = _f32(params), _f32(feat), _f32(labels)
    grad = np.zeros(mlp_num_params(), np.float64) if want_grad else None
    pd = None if params_d is None else np.ascontiguousarray(params_d, dtype=np.float64)
    L = lib()
    L.oracle_mlp_loss_grad.restype = ctypes.c_double
    loss = L.oracle_mlp_loss_grad(_p(params), None if pd is None else _p(pd), _p(feat), _p(labels),
                                  ctypes.c_int64(len(feat)), int(exact), None if grad is None else _p(grad))
    return float(loss), grad


def mlp_adam(params, m, v, grad, t, lr):
    params, m, v, grad = _f32(params).copy(), _f32(m).copy(), _f32(v).copy(), _f32(grad)
    L = lib()
    L.oracle_mlp_adam.restype = None
    L.oracle_mlp_adam(_p(params), _p(m), _p(v), _p(grad), ctypes.c_int64(t), ctypes.c_float(lr))
    return params, m, v


def mlp_gelu(z):
    z = np.ascontiguousarray(z, dtype=np.float64)
    g, dg = np.zeros_like(z), np.zeros_like(z)
    L = lib()
    L.oracle_mlp_gelu.restype = None
    L.oracle_mlp_gelu(_p(z), ctypes.c_int64(len(z)), _p(g), _p(dg))
    return g, dg


def mlp_round_bf16(x):
    x = _f32(x)
    out = np.zeros_like(x)
    L = lib()
    L.oracle_mlp_round_bf16.restype = None
    L.oracle_mlp_round_bf16(_p(x), ctypes.c_int64(len(x)), _p(out))
    return out


# ---- single-function entry points (pinning tests) ---------------------------

def f7_raw(pts, sample):
    """Seven-point models before the orientation filter: array (k, 9)."""
    pts = np.ascontiguousarray(pts, np.float64)
    s = np.ascontiguousarray(sample, np.int32)
    out = np.zeros((3, 9))
    L = lib()
    L.oracle_f7_raw.restype = ctypes.c_int
    k = L.oracle_f7_raw(_p(pts), ctypes.c_int64(pts.shape[0]), _p(s), _p(out))
    return out[:k].copy()


def f_orientation_valid(pts, F, sample):
    pts = np.ascontiguousarray(pts, np.float64)
    F = np.ascontiguousarray(F, np.float64).reshape(9)
    s = np.ascontiguousarray(sample, np.int32)
    L = lib()
    L.oracle_f_orientation_valid.restype = ctypes.c_int
    return bool(L.oracle_f_orientation_valid(_p(pts), ctypes.c_int64(pts.shape[0]), _p(F), _p(s),
                                             ctypes.c_int64(s.size)))


def f_residuals(pts, F):
    """(squared Sampson, squared symmetric epipolar) distance per row."""
    pts = np.ascontiguousarray(pts, np.float64)
    F = np.ascontiguousarray(F, np.float64).reshape(9)
    n = pts.shape[0]
    a = np.zeros(n)
    b = np.zeros(n)
    lib().oracle_f_residuals(_p(pts), ctypes.c_int64(n), _p(F), _p(a), _p(b))
    return a, b


def f8_raw(rows, weights=None):
    rows = np.ascontiguousarray(rows, np.float64)
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    out = np.zeros(9)
    lib().oracle_f8_raw(_p(rows), ctypes.c_int64(rows.shape[0]), _p(w), _p(out))
    return out


def lm_f_normal_equations(rows, F, weights=None):
    """(U, V, sigma, Jtr[7], JtJ[7,7]) of the LM's first step from F."""
    rows = np.ascontiguousarray(rows, np.float64)
    F = np.ascontiguousarray(F, np.float64).reshape(9)
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    U = np.zeros(9)
    V = np.zeros(9)
    sig = ctypes.c_double(0.0)
    Jtr = np.zeros(7)
    JtJ = np.zeros(49)
    lib().oracle_lm_f_normal_equations(_p(rows), ctypes.c_int64(rows.shape[0]), _p(w), _p(F), _p(U), _p(V),
                                       ctypes.byref(sig), _p(Jtr), _p(JtJ))
    return U.reshape(3, 3), V.reshape(3, 3), sig.value, Jtr, JtJ.reshape(7, 7)


def lm_f_refine(rows, F, weights=None, max_iterations=25):
    """(refined F (3, 3), costs: the start then every trial step, accepted
    flag per trial step)."""
    rows = np.ascontiguousarray(rows, np.float64)
    Fio = np.ascontiguousarray(F, np.float64).reshape(9).copy()
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    tr = np.zeros(512)
    acc = np.zeros(512, np.uint8)
    L = lib()
    L.oracle_lm_f_refine.restype = ctypes.c_int
    k = L.oracle_lm_f_refine(_p(rows), ctypes.c_int64(rows.shape[0]), _p(w), _p(Fio), ctypes.c_int32(max_iterations),
                             _p(tr), _p(acc), ctypes.c_int32(512))
    k = min(k, 512)
    return Fio.reshape(3, 3), tr[:k].copy(), acc[:max(k - 1, 0)].astype(bool)


def lm_f_cost(rows, F, weights=None):
    rows = np.ascontiguousarray(rows, np.float64)
    F = np.ascontiguousarray(F, np.float64).reshape(9)
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    L = lib()
    L.oracle_lm_f_cost.restype = ctypes.c_double
    return L.oracle_lm_f_cost(_p(rows), ctypes.c_int64(rows.shape[0]), _p(w), _p(F))


def epipolar_is_valid_model(kind, pts, model, inliers, threshold):
    pts = np.ascontiguousarray(pts, np.float64)
    m = np.ascontiguousarray(model, np.float64).reshape(9)
    inl = np.ascontiguousarray(inliers, np.int64)
    L = lib()
    L.oracle_epipolar_is_valid_model.restype = ctypes.c_int
    return bool(L.oracle_epipolar_is_valid_model(ctypes.c_int32(kind), _p(pts), ctypes.c_int64(pts.shape[0]), _p(m),
                                                 _p(inl), ctypes.c_int64(inl.size), ctypes.c_double(threshold)))


def h_is_valid_sample(pts, sample):
    pts = np.ascontiguousarray(pts, np.float64)
    s = np.ascontiguousarray(sample, np.int32)
    L = lib()
    L.oracle_h_is_valid_sample.restype = ctypes.c_int
    return bool(L.oracle_h_is_valid_sample(_p(pts), ctypes.c_int64(pts.shape[0]), _p(s)))


def h_is_valid_model(H):
    H = np.ascontiguousarray(H, np.float64).reshape(9)
    L = lib()
    L.oracle_h_is_valid_model.restype = ctypes.c_int
    return bool(L.oracle_h_is_valid_model(_p(H)))


def det(a):
    a = np.ascontiguousarray(a, np.float64)
    L = lib()
    L.oracle_det.restype = ctypes.c_double
    return L.oracle_det(_p(a), ctypes.c_int32(a.shape[0]))


def f_post_ba(pts, threshold, model, inliers, score):
    """The driver's post-hoc adjustment: (model (3, 3), inliers, score)."""
    pts = np.ascontiguousarray(pts, np.float64)
    m = np.ascontiguousarray(model, np.float64).reshape(9).copy()
    inl = np.zeros(pts.shape[0], np.int64)
    inl[:len(inliers)] = inliers
    sc = ctypes.c_double(score)
    L = lib()
    L.oracle_f_post_ba.restype = ctypes.c_int64
    k = L.oracle_f_post_ba(_p(pts), ctypes.c_int64(pts.shape[0]), ctypes.c_double(threshold), _p(m), _p(inl),
                           ctypes.c_int64(len(inliers)), ctypes.byref(sc))
    return m.reshape(3, 3), inl[:k].copy(), sc.value


def sprt_sequence(kind, pts, models, threshold, min_ratio, sample_size, max_solutions, seed, best_values):
    """SPRT verifications of models in order (see oracle_sprt_sequence)."""
    pts = np.ascontiguousarray(pts, np.float64)
    models = np.ascontiguousarray(models, np.float64).reshape(-1, 9)
    best = np.ascontiguousarray(best_values, np.float64)
    n, k = pts.shape[0], models.shape[0]
    out = dict(pool=np.zeros(n, np.int64), start=np.zeros(k, np.int64), tested=np.zeros(k, np.int64),
               count=np.zeros(k, np.int64), value=np.zeros(k), accepted=np.zeros(k, np.uint8), eps=np.zeros(k),
               delta=np.zeros(k), A=np.zeros(k))
    lib().oracle_sprt_sequence(
        ctypes.c_int32(kind), _p(pts), ctypes.c_int64(n), _p(models), ctypes.c_int64(k), ctypes.c_double(threshold),
        ctypes.c_double(min_ratio), ctypes.c_int32(sample_size), ctypes.c_int32(max_solutions), ctypes.c_uint32(seed),
        _p(best), _p(out["pool"]), _p(out["start"]), _p(out["tested"]), _p(out["count"]), _p(out["value"]),
        _p(out["accepted"]), _p(out["eps"]), _p(out["delta"]), _p(out["A"]))
    out["accepted"] = out["accepted"].astype(bool)
    return out


def sprt_threshold_A(epsilon, delta, sample_size, max_solutions):
    L = lib()
    L.oracle_sprt_threshold_A.restype = ctypes.c_double
    return float(L.oracle_sprt_threshold_A(ctypes.c_double(epsilon), ctypes.c_double(delta),
                                           ctypes.c_int32(sample_size), ctypes.c_int32(max_solutions)))


def sp_select(pts, cell_sizes, divisions, model, threshold):
    """Space partitioning selection of one homography: per filled cell in the
    grid's iteration order (key, points, axis indices, kept), kept points."""
    pts = np.ascontiguousarray(pts, np.float64)
    cs = np.ascontiguousarray(cell_sizes, np.float64)
    M = np.ascontiguousarray(model, np.float64).ravel()
    n = pts.shape[0]
    keys = np.zeros(n, np.uint64)
    npts = np.zeros(n, np.int64)
    idx = np.zeros(4 * n, np.int64)
    kept = np.zeros(n, np.uint8)
    kp = ctypes.c_int64(0)
    L = lib()
    L.oracle_sp_select.restype = ctypes.c_int64
    c = L.oracle_sp_select(_p(pts), ctypes.c_int64(n), _p(cs), ctypes.c_int64(divisions), _p(M),
                           ctypes.c_double(threshold), ctypes.c_int64(n), _p(keys), _p(npts), _p(idx), _p(kept),
                           ctypes.byref(kp))
    return dict(keys=keys[:c], npts=npts[:c], idx=idx[:4 * c].reshape(c, 4), kept=kept[:c].astype(bool),
                kept_points=kp.value)
