"""Pins the oracle's fundamental-matrix family (and the homography sample /
determinant tests and the essential validity ratio) against the mathematics
that fixes each function, independently of the oracle's own formulas.

Each test names the planted mistake it is built to catch.  CPU only.

Reference functions (src/pygcransac/include/...):
  seven-point      estimators/solver_fundamental_matrix_seven_point.h:91-291
  orientation      estimators/fundamental_estimator.h:737-800
  Sampson          estimators/fundamental_estimator.h:195-222
  symmetric        estimators/fundamental_estimator.h:224-246
  eight-point      estimators/solver_fundamental_matrix_eight_point.h:93-182
  LM               relative_pose/bundle.cpp:252-339, relative_pose/jacobian_impl.h:468-614
  F validity       estimators/fundamental_estimator.h:287-323
  E validity       estimators/essential_estimator.h:279-305
  H sample test    estimators/homography_estimator.h:346-381
  H model test     estimators/homography_estimator.h:326-342
  post-hoc BA      src/gcransac_python.cpp:1832-1887
"""
from fractions import Fraction

import numpy as np
import pytest


# ---- independent geometry (numpy / exact rationals) -------------------------

def _rodrigues(w):
    th = np.linalg.norm(w)
    if th == 0.0:
        return np.eye(3)
    k = w / th
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K


def _two_view(rng, n, behind_second=0):
    """Noise-free correspondences of n points in front of both cameras, in
    normalised coordinates (f = 1).  With behind_second > 0 the second camera
    sits 3 units ahead of the first and that many more points lie between
    them: in front of camera 1, behind camera 2.  Returns corr (n', 4), F."""
    R = _rodrigues(rng.normal(size=3) * 0.2)
    if behind_second:
        C2 = np.array([rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), 3.0])
        t = -R @ C2
    else:
        t = rng.normal(size=3)
        t /= np.linalg.norm(t)
    X = np.c_[rng.uniform(-1, 1, (n, 2)) * 3, rng.uniform(4, 10, n)]
    if behind_second:
        X = np.c_[rng.uniform(-1, 1, (n, 2)), rng.uniform(5, 10, n)]
    X2 = X @ R.T + t
    assert (X2[:, 2] > 0).all()
    rows = [np.c_[X[:, :2] / X[:, 2:], X2[:, :2] / X2[:, 2:]]]
    if behind_second:
        Xb = np.c_[rng.uniform(-0.3, 0.3, (behind_second, 2)), rng.uniform(1.0, 2.0, behind_second)]
        Xb2 = Xb @ R.T + t
        assert (Xb2[:, 2] < 0).all()
        rows.append(np.c_[Xb[:, :2] / Xb[:, 2:], Xb2[:, :2] / Xb2[:, 2:]])
    tx = np.array([[0, -t[2], t[1]], [t[2], 0, -t[0]], [-t[1], t[0], 0]])
    F = tx @ R
    return np.vstack(rows), F / np.linalg.norm(F)


def _h(p):
    return np.c_[p, np.ones(len(p))]


def _epi(corr, F):
    """x2^T F x1 per row."""
    return np.einsum("ij,jk,ik->i", _h(corr[:, 2:4]), F, _h(corr[:, 0:2]))


def _sampson_fd(corr, F):
    """Sampson distance by definition: C^2 / |grad C|^2, C = x2^T F x1 as a
    function of (x1, y1, x2, y2), the gradient by central differences (C is
    bilinear, so the difference quotient is exact up to rounding)."""
    h = 1e-3
    C = _epi(corr, F)
    g2 = np.zeros(len(corr))
    for j in range(4):
        e = np.zeros(4)
        e[j] = h
        g2 += ((_epi(corr + e, F) - _epi(corr - e, F)) / (2 * h)) ** 2
    return C * C / g2


def _symmetric_lines(corr, F):
    """Squared distance of x2 to the line F x1 plus of x1 to the line F^T x2."""
    l2 = _h(corr[:, 0:2]) @ F.T          # lines in image 2
    l1 = _h(corr[:, 2:4]) @ F            # lines in image 1
    d2 = (np.sum(l2 * _h(corr[:, 2:4]), 1)) ** 2 / (l2[:, 0] ** 2 + l2[:, 1] ** 2)
    d1 = (np.sum(l1 * _h(corr[:, 0:2]), 1)) ** 2 / (l1[:, 0] ** 2 + l1[:, 1] ** 2)
    return d1 + d2


def _unit(F):
    F = np.asarray(F, np.float64).reshape(3, 3)
    F = F / np.linalg.norm(F)
    i = np.argmax(np.abs(F))
    return F * np.sign(F.flat[i])


# ---- seven-point solver ---------------------------------------------------------

def _pencil_models(corr7):
    """Rank-2 members of the 2-D null space of the 7 x 9 system
    x2^T F x1 = 0 (numpy SVD basis, cubic by interpolation, numpy.roots)."""
    A = np.array([np.kron(_h(corr7[i:i + 1, 2:4])[0], _h(corr7[i:i + 1, 0:2])[0]) for i in range(7)])
    _, s, Vt = np.linalg.svd(A)
    N1, N2 = Vt[-1].reshape(3, 3), Vt[-2].reshape(3, 3)
    xs = np.array([-2.0, -1.0, 0.5, 1.5])
    ys = np.array([np.linalg.det(x * N1 + N2) for x in xs])
    coef = np.polyfit(xs, ys, 3)
    roots = np.roots(coef)
    out, ambiguous = [], False
    for r in roots:
        if abs(r.imag) < 1e-9 * max(1.0, abs(r)):
            out.append(_unit(r.real * N1 + N2))
        elif abs(r.imag) < 1e-3 * max(1.0, abs(r)):
            ambiguous = True
    if abs(np.linalg.det(N1)) < 1e-12:
        ambiguous = True  # a root at infinity
    return out, ambiguous


def test_seven_point_models_are_the_rank2_members_of_the_pencil(oracle):
    """Catches: a wrong determinant-cubic coefficient (models with det F != 0),
    a transposed system row (x2^T F x1 != 0 on the sample), a dropped real
    root or a spurious one (count differs from numpy.roots of the cubic)."""
    rng = np.random.default_rng(7)
    checked = 0
    for trial in range(300):
        if trial % 2 == 0:
            corr, _ = _two_view(rng, 7)
            corr = corr + rng.normal(0, 1e-3, corr.shape)   # generic, near-consistent
        else:
            corr = rng.uniform(-1, 1, (7, 4))                # arbitrary 7 matches
        ms = oracle.f7_raw(corr, np.arange(7))
        ref, ambiguous = _pencil_models(corr)
        for m in ms:
            F = m.reshape(3, 3)
            nf = np.linalg.norm(F)
            assert abs(np.linalg.det(F)) < 1e-9 * nf ** 3
            x1, x2 = _h(corr[:, :2]), _h(corr[:, 2:])
            res = np.abs(_epi(corr, F)) / (nf * np.linalg.norm(x1, axis=1) * np.linalg.norm(x2, axis=1))
            assert res.max() < 1e-9
        if ambiguous or len(ms) == 0:
            continue
        assert len(ms) == len(ref), (trial, len(ms), len(ref))
        got = [_unit(m) for m in ms]
        for R in ref:
            assert min(np.linalg.norm(R - G) for G in got) < 1e-6
        checked += 1
    assert checked > 250


def test_seven_point_recovers_the_true_fundamental_matrix(oracle):
    """Catches a solver that returns a rank-2 member of the wrong pencil."""
    rng = np.random.default_rng(11)
    for _ in range(50):
        corr, F = _two_view(rng, 7)
        ms = oracle.f7_raw(corr, np.arange(7))
        assert min(np.linalg.norm(_unit(m) - _unit(F)) for m in ms) < 1e-7


def test_orientation_filter_accepts_points_in_front_and_rejects_one_behind(oracle):
    """Catches a wrong epipole (row cross product) or sign test: points in
    front of both cameras satisfy the oriented epipolar constraint for F and
    -F; a sample with one point behind the second camera does not."""
    rng = np.random.default_rng(3)
    for _ in range(100):
        corr, F = _two_view(rng, 6, behind_second=1)
        front = np.arange(6)
        assert oracle.f_orientation_valid(corr, F, front)
        assert oracle.f_orientation_valid(corr, -F, front)
        assert not oracle.f_orientation_valid(corr, F, np.arange(7))
        assert not oracle.f_orientation_valid(corr, -F, np.r_[6, front])


# ---- Sampson and symmetric epipolar distances -----------------------------------

def _exact_sampson_symmetric(pt, F):
    x1, y1, x2, y2 = (Fraction(v) for v in pt)
    Fr = [[Fraction(v) for v in r] for r in F]
    X1, X2 = [x1, y1, Fraction(1)], [x2, y2, Fraction(1)]
    Fx1 = [sum(Fr[i][j] * X1[j] for j in range(3)) for i in range(3)]
    Ftx2 = [sum(Fr[i][j] * X2[i] for i in range(3)) for j in range(3)]
    C = sum(X2[i] * Fx1[i] for i in range(3))
    # dC/dx1 = (F^T x2)_0, dC/dy1 = (F^T x2)_1, dC/dx2 = (F x1)_0, dC/dy2 = (F x1)_1
    a = Ftx2[0] ** 2 + Ftx2[1] ** 2
    b = Fx1[0] ** 2 + Fx1[1] ** 2
    sampson = C * C / (a + b)
    symmetric = C * C / a + C * C / b
    return float(sampson), float(symmetric)


def test_sampson_and_symmetric_exact_rationals(oracle):
    """Small integer inputs evaluated exactly with fractions: catches a dropped
    or swapped denominator term (e.g. (F x2) in place of (F^T x2)), which the
    exact value exposes by far more than rounding."""
    rng = np.random.default_rng(5)
    for _ in range(200):
        F = rng.integers(-9, 10, (3, 3)).astype(np.float64)
        pts = rng.integers(-20, 21, (6, 4)).astype(np.float64)
        samp, sym = oracle.f_residuals(pts, F)
        for i in range(6):
            try:
                es, ey = _exact_sampson_symmetric(pts[i], F)
            except ZeroDivisionError:
                continue
            assert samp[i] == pytest.approx(es, rel=1e-14, abs=1e-300)
            assert sym[i] == pytest.approx(ey, rel=1e-14, abs=1e-300)


def test_sampson_is_first_order_error_and_symmetric_is_line_distance(oracle):
    """Sampson = C^2 / |grad C|^2 with the gradient by central differences;
    symmetric = sum of squared point-to-epipolar-line distances in both
    images; symmetric >= Sampson.  Pixel-sized inputs."""
    rng = np.random.default_rng(9)
    corr, F = _two_view(rng, 400)
    K = np.array([[800, 0, 320], [0, 800, 240], [0, 0, 1.0]])
    px = np.c_[corr[:, :2] * 800 + [320, 240], corr[:, 2:] * 800 + [320, 240]]
    px = px + rng.normal(0, 2.0, px.shape)
    Fp = np.linalg.inv(K).T @ F @ np.linalg.inv(K)
    samp, sym = oracle.f_residuals(px, Fp)
    np.testing.assert_allclose(samp, _sampson_fd(px, Fp), rtol=1e-7)
    np.testing.assert_allclose(sym, _symmetric_lines(px, Fp), rtol=1e-9)
    assert (sym >= samp * (1 - 1e-12)).all()


# ---- eight-point solver -----------------------------------------------------------

def _design(rows, w=None):
    A = np.array([np.kron(_h(rows[i:i + 1, 2:4])[0], _h(rows[i:i + 1, 0:2])[0]) for i in range(len(rows))])
    return A if w is None else A * w[:, None]


def test_eight_point_exact_recovery_weighted_and_unweighted(oracle):
    """Noise-free rows: the null vector is the true F, with or without
    weights.  Catches a transposed row layout or a wrong coefficient."""
    rng = np.random.default_rng(21)
    for _ in range(40):
        corr, F = _two_view(rng, 30)
        assert np.linalg.norm(_unit(oracle.f8_raw(corr)) - _unit(F)) < 1e-9
        w = rng.uniform(0.1, 1.0, len(corr))
        assert np.linalg.norm(_unit(oracle.f8_raw(corr, w)) - _unit(F)) < 1e-9


def test_eight_point_vector_is_orthogonal_to_eight_columns_of_weighted_AtA(oracle):
    """The solver takes the last column q of Q in a full-pivot QR of
    M = (W A)^T (W A); q is orthogonal to the eight pivot columns, so M q has
    eight zero entries.  With noisy rows and uneven weights this fails if the
    weight enters as w^2 per coefficient (or not at all).  q is also close to
    the least-squares null vector: |W A q| / |q| within 1.25 times the
    smallest singular value (numpy)."""
    rng = np.random.default_rng(22)
    for _ in range(40):
        corr, F = _two_view(rng, 60)
        corr = corr + rng.normal(0, 2e-3, corr.shape)
        w = rng.uniform(0.05, 1.0, len(corr))
        q = oracle.f8_raw(corr, w)
        A = _design(corr, w)
        M = A.T @ A
        v = np.sort(np.abs(M @ q))
        assert v[7] < 1e-12 * np.abs(M).max(), v
        smin = np.linalg.svd(A, compute_uv=False)[-1]
        assert np.linalg.norm(A @ q) / np.linalg.norm(q) < 1.25 * smin


# ---- Levenberg-Marquardt --------------------------------------------------------

def _factor_F(U, V, sigma, p):
    Up = _rodrigues(p[0:3]) @ U
    Vp = _rodrigues(p[3:6]) @ V
    return np.outer(Up[:, 0], Vp[:, 0]) + (sigma + p[6]) * np.outer(Up[:, 1], Vp[:, 1])


def _signed_sampson(rows, F):
    x1, x2 = _h(rows[:, :2]), _h(rows[:, 2:])
    C = np.sum(x2 * (x1 @ F.T), 1)
    Fx1 = x1 @ F.T
    Ftx2 = x2 @ F
    return C / np.sqrt(Ftx2[:, 0] ** 2 + Ftx2[:, 1] ** 2 + Fx1[:, 0] ** 2 + Fx1[:, 1] ** 2)


def test_lm_normal_equations_match_finite_differences(oracle):
    """Jtr and JtJ of the LM step against central differences of the signed
    Sampson residual along the step's own parametrisation (U <- R(w1) U,
    V <- R(w2) V, sigma <- sigma + d).  Catches a sign error or dropped term
    in the Sampson Jacobian or the dF/dparams table, and a wrong weight."""
    rng = np.random.default_rng(31)
    for trial in range(12):
        corr, F = _two_view(rng, 80)
        rows = corr + rng.normal(0, 2e-3, corr.shape)
        w = rng.uniform(0.2, 1.0, len(rows)) if trial % 2 else None
        # start from a rank-2 perturbation of the truth
        U, s, Vt = np.linalg.svd(F)
        Fs = (_rodrigues(rng.normal(size=3) * 0.01) @ U) @ np.diag([1.0, s[1] / s[0] * 1.01, 0.0]) @ \
            (_rodrigues(rng.normal(size=3) * 0.01) @ Vt.T).T
        Uo, Vo, sig, Jtr, JtJ = oracle.lm_f_normal_equations(rows, Fs, w)
        # the factorisation reproduces F up to scale
        Frec = _factor_F(Uo, Vo, sig, np.zeros(7))
        assert np.linalg.norm(_unit(Frec) - _unit(Fs)) < 1e-10
        r0 = _signed_sampson(rows, Frec)
        assert (r0 ** 2 < 0.5).all()  # away from the truncation step
        n = len(rows)
        wk = np.full(n, 1.0 / n) * (1.0 if w is None else w)
        h = 1e-6
        J = np.zeros((n, 7))
        for j in range(7):
            e = np.zeros(7)
            e[j] = h
            J[:, j] = (_signed_sampson(rows, _factor_F(Uo, Vo, sig, e)) -
                       _signed_sampson(rows, _factor_F(Uo, Vo, sig, -e))) / (2 * h)
        g = J.T @ (wk * r0)
        H = (J * wk[:, None]).T @ J
        assert np.linalg.norm(Jtr - g) < 1e-6 * np.linalg.norm(g) + 1e-15
        assert np.linalg.norm(JtJ - H) < 1e-6 * np.linalg.norm(H)


def _accepted_costs(trace, acc):
    """Replays the acceptance rule over the trial costs: every kept step
    lowered the cost, every rejected one did not.  Returns the kept costs and
    the number of rejections."""
    cur = trace[0]
    kept = [cur]
    for c, a in zip(trace[1:], acc):
        if a:
            assert c < cur
            cur = c
            kept.append(c)
        else:
            assert not (c < cur)
    return np.array(kept), int((~acc).sum())


def test_lm_keeps_only_improving_steps_and_recovers_noise_free_truth(oracle):
    """Kept steps lower the truncated cost and rejected ones would not have
    (replayed over every trial step, from starts far enough away that some
    steps are rejected); on noise-free rows a perturbed start converges to the
    true F.  Catches a step applied with the wrong sign or an acceptance test
    that keeps worse steps."""
    rng = np.random.default_rng(41)
    rejections = 0
    for trial in range(40):
        corr, F = _two_view(rng, 60)
        U, s, Vt = np.linalg.svd(F)
        a = 0.02 if trial % 2 == 0 else 0.25
        if trial % 2:
            corr = corr + rng.normal(0, 3e-3, corr.shape)  # noisy rows, far start: rejected steps
        Fs = (_rodrigues(rng.normal(size=3) * a) @ U) @ np.diag([1.0, s[1] / s[0] * (1 + a), 0.0]) @ \
            (_rodrigues(rng.normal(size=3) * a) @ Vt.T).T
        Fr, trace, acc = oracle.lm_f_refine(corr, Fs)
        kept, rej = _accepted_costs(trace, acc)
        rejections += rej
        assert oracle.lm_f_cost(corr, Fr) == kept[-1]
        if trial % 2 == 0:
            # the LM stops once its step or gradient norm is below 1e-8
            assert kept[-1] < 1e-8 * kept[0]
            assert np.linalg.norm(_unit(Fr) - _unit(F)) < 1e-4
    assert rejections >= 5


def test_lm_result_is_a_stationary_point_of_the_cost(oracle):
    """On noisy rows the refined F zeroes the cost gradient along the rank-2
    manifold (central differences of the numpy cost), from a start where the
    gradient is large."""
    rng = np.random.default_rng(43)
    for trial in range(8):
        corr, F = _two_view(rng, 100)
        rows = corr + rng.normal(0, 3e-3, corr.shape)
        U, s, Vt = np.linalg.svd(F)
        Fs = (_rodrigues(rng.normal(size=3) * 0.03) @ U) @ np.diag([1.0, s[1] / s[0], 0.0]) @ Vt
        Fr, trace, acc = oracle.lm_f_refine(rows, Fs, max_iterations=200)

        def grad(Fm):
            Uo, Vo, sig, _, _ = oracle.lm_f_normal_equations(rows, Fm)
            h = 1e-6
            g = np.zeros(7)
            for j in range(7):
                e = np.zeros(7)
                e[j] = h
                cp = np.mean(np.minimum(_signed_sampson(rows, _factor_F(Uo, Vo, sig, e)) ** 2, 1.0))
                cm = np.mean(np.minimum(_signed_sampson(rows, _factor_F(Uo, Vo, sig, -e)) ** 2, 1.0))
                g[j] = (cp - cm) / (2 * h)
            return g
        g0, g1 = grad(Fs), grad(Fr)
        assert np.linalg.norm(g1) < 1e-3 * np.linalg.norm(g0)
        assert oracle.lm_f_cost(rows, Fr) <= oracle.lm_f_cost(rows, Fs)


# ---- validity tests ---------------------------------------------------------------

def _graded_points(corr, F, n_exact, n_mid, n_far):
    """From noise-free rows: copies whose symmetric distance (line-distance
    definition) is ~0 (exact), in (1.5, 3.5) with Sampson below 1 (mid), or
    above 10 (far), by moving x2 along the normal of its epipolar line."""
    exact, mid, far = [], [], []
    for row in corr:
        l2 = F @ np.r_[row[0:2], 1.0]
        nrm = l2[:2] / np.hypot(l2[0], l2[1])
        for target, out in ((0.0, exact), (1.0, mid), (15.0, far)):
            p = row.copy()
            p[2:4] += target * nrm
            sym = _symmetric_lines(p[None], F)[0]
            samp = _sampson_fd(p[None], F)[0]
            if target == 0.0 and sym < 1e-12:
                out.append(p)
            elif target == 1.0 and 1.5 < sym < 3.5 and samp < 1.0:
                out.append(p)
            elif target == 15.0 and sym > 10.0:
                out.append(p)
    assert len(exact) >= n_exact and len(mid) >= n_mid and len(far) >= n_far, (len(exact), len(mid), len(far))
    return np.array(exact[:n_exact]), np.array(mid[:n_mid]), np.array(far[:n_far])


@pytest.mark.parametrize("kind,ratio,m", [(1, 0.5, 7), (2, 0.1, 5)])
def test_epipolar_validity_ratio_counts_symmetric_inliers(oracle, kind, ratio, m):
    """A model is valid when at least max(m, ratio * |I|) (truncated) of the
    inliers are also inliers under the symmetric epipolar distance.  The
    'mid' points pass Sampson but fail the symmetric test, so they catch a
    validity test that uses the Sampson distance; the boundary cases catch
    the wrong ratio, m, rounding or > in place of >=."""
    rng = np.random.default_rng(50 + kind)
    corr, F = _two_view(rng, 600)
    K = np.array([[800, 0, 320], [0, 800, 240], [0, 0, 1.0]])
    px = np.c_[corr[:, :2] * 800 + [320, 240], corr[:, 2:] * 800 + [320, 240]]
    Fp = np.linalg.inv(K).T @ F @ np.linalg.inv(K)
    ex, mid, far = _graded_points(px, Fp, 150, 80, 80)
    F = Fp
    pts = np.vstack([ex, mid, far])
    ie = np.arange(len(ex))
    im = len(ex) + np.arange(len(mid))
    iff = len(ex) + len(mid) + np.arange(len(far))
    for total in (10, 20, 21, 40, 57, 100, 150):
        need = int(max(m, total * ratio))
        for good in (need - 1, need):
            if good < 0 or good > total:
                continue
            bad = total - good
            inl = np.r_[ie[:good], im[:bad // 2], iff[:bad - bad // 2]]
            rng.shuffle(inl)
            assert oracle.epipolar_is_valid_model(kind, pts, F, inl, 1.0) == (good >= need), (total, good, need)


def test_homography_sample_test_is_orientation_consistency(oracle):
    """Samples related by an orientation-preserving homography (det H > 0,
    points in front) pass; mirrored samples (det H < 0) fail; arbitrary
    samples pass iff the four triangle orientations agree between images.
    Catches a swapped cross product or a flipped comparison."""
    rng = np.random.default_rng(61)

    def orient(a, b, c):
        return np.linalg.det(np.array([[a[0], a[1], 1.0], [b[0], b[1], 1.0], [c[0], c[1], 1.0]]))

    for trial in range(300):
        x1 = rng.uniform(0, 640, (4, 2))
        th = rng.uniform(-0.5, 0.5)
        A = np.array([[np.cos(th), -np.sin(th), rng.uniform(-40, 40)],
                      [np.sin(th), np.cos(th), rng.uniform(-40, 40)],
                      [rng.normal(0, 1e-4), rng.normal(0, 1e-4), 1.0]])
        X = _h(x1) @ A.T
        x2 = X[:, :2] / X[:, 2:]
        pts = np.c_[x1, x2]
        assert oracle.h_is_valid_sample(pts, np.arange(4))
        mir = pts.copy()
        mir[:, 2] = 640 - mir[:, 2]
        tri = [(0, 1, 2), (0, 1, 3), (2, 3, 0), (2, 3, 1)]
        if min(abs(orient(*x1[list(t)])) for t in tri) > 1e-3:
            assert not oracle.h_is_valid_sample(mir, np.arange(4))
        rnd = rng.uniform(0, 640, (4, 4))
        agree = all(orient(*rnd[list(t), 0:2]) * orient(*rnd[list(t), 2:4]) >= 0 for t in tri)
        assert oracle.h_is_valid_sample(rnd, np.arange(4)) == agree


def test_homography_model_test_is_determinant_magnitude(oracle):
    """|det H| >= 1e-2 (homography_estimator.h:326-342); the determinant
    agrees with numpy's.  Catches a signed comparison or another bound."""
    rng = np.random.default_rng(71)
    for _ in range(200):
        M = rng.normal(size=(3, 3))
        assert oracle.det(M) == pytest.approx(np.linalg.det(M), rel=1e-12, abs=1e-15)
    for d in (0.5e-2, 0.99e-2, 1.01e-2, 2e-2, 1.0):
        for sgn in (1.0, -1.0):
            H = np.diag([1.0, 1.0, sgn * d])
            assert oracle.h_is_valid_model(H) == (d >= 1e-2)


# ---- post-hoc bundle adjustment ----------------------------------------------------

def _msac(pts, F, sq):
    r2 = _sampson_fd(pts, F)
    keep = r2 < sq
    if not keep.any():
        return 0.0, np.zeros(0, np.int64)
    v = -np.sum(r2[keep])
    return (v + keep.sum() * sq) / sq, np.nonzero(keep)[0]


def test_post_hoc_bundle_adjustment_weights_and_replacement_rule(oracle):
    """The candidate is the weighted bundle adjustment with weights
    max(0, 1 - r^2 / (1.5 t)^2) by inlier position; it replaces the model
    only when its MSAC value at 1.5 t is higher.  Catches weights at t
    instead of 1.5 t, and a rule that keeps a worse candidate."""
    rng = np.random.default_rng(81)
    K = np.array([[800, 0, 320], [0, 800, 240], [0, 0, 1.0]])
    Ki = np.linalg.inv(K)
    replaced = kept = 0
    for trial in range(32):
        corr, F = _two_view(rng, 150)
        px = np.c_[corr[:, :2] * 800 + [320, 240], corr[:, 2:] * 800 + [320, 240]]
        px = px + rng.normal(0, 0.5, px.shape)
        out = rng.uniform(0, 640, (100, 4))
        pts = np.vstack([px, out])
        # the true model, or a slightly rotated one (three trials in four)
        U, s, Vt = np.linalg.svd(F)
        a = 0.0 if trial % 4 < 2 else 1e-3
        Fn = (_rodrigues(rng.normal(size=3) * a) @ U) @ np.diag([1.0, s[1] / s[0], 0.0]) @ \
            (_rodrigues(rng.normal(size=3) * a) @ Vt.T).T
        Fs = Ki.T @ Fn @ Ki
        t = 3.0 if trial % 2 else 1.0
        sq = (1.5 * t) ** 2
        score0, inl0 = _msac(pts, Fs, sq)
        Fo, inl, score = oracle.f_post_ba(pts, t, Fs, inl0, score0)
        if len(inl0) <= 7:
            np.testing.assert_array_equal(Fo, np.eye(3))
            continue
        w = np.maximum(0.0, 1.0 - _sampson_fd(pts[inl0], Fs) / sq)
        cand = oracle.fit_nonminimal(1, pts, inl0, weights=w, weight_mode=1, normalize=0, threshold=t)
        cand2 = oracle.fit_nonminimal(1, pts, inl0, weights=Fs.reshape(9), weight_mode=2, normalize=0, threshold=t)
        np.testing.assert_allclose(cand, cand2, rtol=1e-9, atol=1e-12)
        best = score0
        expect = Fs
        for c in cand:
            s, _ = _msac(pts, c.reshape(3, 3), sq)
            if s > best * (1 + 1e-12):
                best, expect = s, c.reshape(3, 3)
        assert score >= score0 * (1 - 1e-12)
        assert score == pytest.approx(best, rel=1e-9)
        np.testing.assert_allclose(Fo, expect, rtol=1e-9, atol=1e-12)
        if np.array_equal(Fo, Fs):
            kept += 1
        else:
            replaced += 1
            s, inl_py = _msac(pts, Fo, sq)
            np.testing.assert_array_equal(np.sort(inl), inl_py)
    assert replaced > 0 and kept > 0


def test_post_hoc_bundle_adjustment_small_sets_give_identity(oracle):
    """Seven or fewer inliers: the driver returns the identity."""
    rng = np.random.default_rng(83)
    pts = rng.uniform(0, 640, (20, 4))
    Fo, inl, score = oracle.f_post_ba(pts, 1.0, np.ones((3, 3)), np.arange(7), 3.0)
    np.testing.assert_array_equal(Fo, np.eye(3))
