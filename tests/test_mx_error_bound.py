"""The error bound of the matrix-core scoring pass, checked on the CPU.

The kernel (csrc/k_score_mx.hip) drops a point for a model only when the
low-precision quantities prove it an outlier, using
delta = 2^-17 S + 2^-20 for a quantity c . z with |c_t| <= 1, S = sum of |z_t|
over the terms it reads.  This test rebuilds the kernel's operands exactly as
it does -- z in FP64 then rounded to FP32, FP16 hi = fp16(z), lo = fp16(z - hi)
(the coefficients split from FP64) -- forms the 27 products hi*hi, hi*lo,
lo*hi (exact in FP32) and accumulates them in FP32 in several orders (the
MFMA's internal order is not specified), then compares with the exact value
(FP64 from the unsplit operands, whose own error is ~1e-16).  Inputs include
near-threshold cases, tiny and subnormal-range operands, and coordinates at
the scale limit |x'| = 1.  The measured worst case is reported as a fraction
of the bound; the bound must hold everywhere.
"""
import numpy as np


def _split16(v):
    hi = v.astype(np.float16)
    lo = (v - hi.astype(v.dtype)).astype(np.float16)
    return hi, lo


def _z(x1, y1, x2, y2):
    return np.stack([x2 * x1, x2 * y1, x2, y2 * x1, y2 * y1, y2, x1, y1, np.ones_like(x1)], axis=-1)


def _quantities(F):
    """Coefficient vectors of C, (F x1)_0, (F x1)_1, (F^T x2)_0, (F^T x2)_1
    and of the transfer quantities P, Q, W (as the kernel builds them)."""
    n = F.shape[0]
    c = np.zeros((n, 8, 9))
    c[:, 0] = F
    c[:, 1, 6:9] = F[:, 0:3]
    c[:, 2, 6:9] = F[:, 3:6]
    c[:, 3, [2, 5, 8]] = F[:, [0, 3, 6]]
    c[:, 4, [2, 5, 8]] = F[:, [1, 4, 7]]
    c[:, 5, [0, 1, 2]] = F[:, 6:9]
    c[:, 5, [6, 7, 8]] = -F[:, 0:3]
    c[:, 6, [3, 4, 5]] = F[:, 6:9]
    c[:, 6, [6, 7, 8]] = -F[:, 3:6]
    c[:, 7, [6, 7, 8]] = F[:, 6:9]
    return c


def _sum_terms(q):
    """Which z terms each quantity reads -> the S of its bound."""
    return [[0, 1, 2, 3, 4, 5, 6, 7, 8], [6, 7, 8], [6, 7, 8], [2, 5, 8], [2, 5, 8],
            [0, 1, 2, 6, 7, 8], [3, 4, 5, 6, 7, 8], [6, 7, 8]][q]


def _bound_sets():
    # the kernel's three bounds: d[0] uses (1+|x1|+|y1|)(1+|x2|+|y2|) (C, P, Q),
    # d[1] 1+|x1|+|y1| ((F x1), W), d[2] 1+|x2|+|y2| ((F^T x2))
    return [0, 1, 1, 2, 2, 0, 0, 1]


def test_low_precision_error_within_kernel_bound():
    rng = np.random.default_rng(17)
    worst = 0.0
    for trial in range(40):
        nm, npts = 64, 256
        # models with |F''| <= 1: random, scaled, with tiny and zero entries
        F = rng.uniform(-1, 1, (nm, 9))
        F *= 10.0 ** rng.uniform(-7, 0, (nm, 9))
        F[rng.uniform(size=F.shape) < 0.05] = 0.0
        F /= np.abs(F).max(axis=1, keepdims=True)
        # scaled points |x'| <= 1, some at the limit, some tiny
        P = rng.uniform(-1, 1, (npts, 4)) * (10.0 ** rng.uniform(-6, 0, (npts, 4)))
        P[: npts // 8] = np.sign(P[: npts // 8]) * 1.0
        x1, y1, x2, y2 = P.T
        zd = _z(x1, y1, x2, y2)                      # FP64
        z32 = zd.astype(np.float32)
        zh, zl = _split16(z32)
        c = _quantities(F)                           # (nm, 8, 9) FP64
        ch, cl = _split16(c)
        exact = np.einsum("mqt,pt->mqp", c, zd)      # FP64 reference
        # the 27 products, exact in FP32 (11-bit x 11-bit significands)
        prods = np.concatenate([
            np.einsum("mqt,pt->mqpt", ch.astype(np.float32), zh.astype(np.float32)),
            np.einsum("mqt,pt->mqpt", ch.astype(np.float32), zl.astype(np.float32)),
            np.einsum("mqt,pt->mqpt", cl.astype(np.float32), zh.astype(np.float32))], axis=-1)
        sa = 1 + np.abs(x1) + np.abs(y1)
        sb = 1 + np.abs(x2) + np.abs(y2)
        S = np.stack([sa * sb, sa, sb])              # (3, npts)
        delta = 2.0 ** -17 * S + 2.0 ** -20
        bsel = _bound_sets()
        for order in range(3):
            perm = rng.permutation(27) if order else np.arange(27)
            acc = np.zeros(prods.shape[:3], np.float32)
            for k in perm:
                acc = (acc + prods[..., k]).astype(np.float32)
            err = np.abs(acc.astype(np.float64) - exact)
            for q in range(8):
                ratio = err[:, q, :] / delta[bsel[q]][None, :]
                worst = max(worst, float(ratio.max()))
                assert (ratio <= 1.0).all(), (trial, order, q, float(ratio.max()))
                # the S the kernel uses covers the terms the quantity reads
                Sq = np.abs(zd[:, _sum_terms(q)]).sum(axis=1)
                assert (Sq <= S[bsel[q]] * (1 + 1e-12)).all()
    print("worst low-precision error: %.3f of the kernel's bound" % worst)
    assert worst < 0.5


def test_rejection_never_drops_an_inlier():
    """End-to-end on scaled Sampson quantities: the kernel's rejection test,
    evaluated on the emulated low-precision values, never fires for a point
    whose exact squared Sampson distance is below the threshold; and it does
    fire for most far points (the pass is useful)."""
    rng = np.random.default_rng(23)
    rejected_far = total_far = 0
    for trial in range(30):
        # a rank-2 F'' and points near and far from its epipolar geometry
        U, _, Vt = np.linalg.svd(rng.normal(size=(3, 3)))
        F = (U @ np.diag([1.0, rng.uniform(0.2, 1.0), 0.0]) @ Vt).ravel()
        F /= np.abs(F).max()
        n = 2000
        x1 = rng.uniform(-1, 1, (n, 2))
        l2 = np.c_[x1, np.ones(n)] @ F.reshape(3, 3).T
        # x2 on its epipolar line, then moved off it by d (normalised units)
        t = rng.uniform(-1, 1, n)
        nrm = l2[:, :2] / np.linalg.norm(l2[:, :2], axis=1, keepdims=True)
        foot = -l2[:, 2:3] * nrm / np.linalg.norm(l2[:, :2], axis=1, keepdims=True)
        x2 = foot + t[:, None] * nrm[:, ::-1] * [1, -1]
        d = 10.0 ** rng.uniform(-5, -1, n) * rng.choice([-1, 1], n)
        x2 = x2 + d[:, None] * nrm
        keep = (np.abs(x2) <= 1).all(axis=1)
        x1, x2 = x1[keep], x2[keep]
        zd = _z(x1[:, 0], x1[:, 1], x2[:, 0], x2[:, 1])
        zh, zl = _split16(zd.astype(np.float32))
        c = _quantities(F[None])[0, :5]
        ch, cl = _split16(c)
        q = (ch.astype(np.float32) @ zh.astype(np.float32).T + ch.astype(np.float32) @ zl.astype(np.float32).T +
             cl.astype(np.float32) @ zh.astype(np.float32).T).astype(np.float32)
        ex = c @ zd.T
        sampson = ex[0] ** 2 / (ex[1] ** 2 + ex[2] ** 2 + ex[3] ** 2 + ex[4] ** 2)
        tau2 = (2e-3) ** 2
        sa = (1 + np.abs(x1).sum(1)).astype(np.float32)
        sb = (1 + np.abs(x2).sum(1)).astype(np.float32)
        dC = np.float32(2.0 ** -17) * sa * sb + np.float32(2.0 ** -20)
        da = np.float32(2.0 ** -17) * sa + np.float32(2.0 ** -20)
        db = np.float32(2.0 ** -17) * sb + np.float32(2.0 ** -20)
        den = ((np.abs(q[1]) + da) ** 2 + (np.abs(q[2]) + da) ** 2 + (np.abs(q[3]) + db) ** 2 +
               (np.abs(q[4]) + db) ** 2).astype(np.float32)
        cl_ = (np.abs(q[0]) - dC).astype(np.float32)
        tau2m = np.float32(tau2 * (1 + 2.0 ** -16))
        rej = (cl_ > 0) & (cl_ * cl_ > den * tau2m)
        assert not (rej & (sampson < tau2)).any(), trial
        far = sampson > 4 * tau2
        rejected_far += int((rej & far).sum())
        total_far += int(far.sum())
    assert rejected_far > 0.95 * total_far
