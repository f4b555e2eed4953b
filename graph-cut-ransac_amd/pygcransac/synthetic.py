"""Seeded synthetic two-view scenes for tests and benchmarks.

Each generator returns a dict with the (N, 4) float64 correspondence array
``corr`` (x1, y1, x2, y2 per row), the boolean ground-truth inlier mask
``gt``, the image sizes and the true model.  Rows are shuffled.  All
randomness comes from ``numpy.random.default_rng(seed)`` so a (kind, seed)
pair always yields the same scene.

Scene families:
  homography  -- plane-induced H between two 640x480 images: rotation
                 +-30 deg, scale 0.7..1.4, perspective terms ~1e-4,
                 translation +-40 px, about the image centre.
  fundamental -- 3-D points in the first camera's frustum at depth 2..20 m,
                 two pinhole cameras (f=800, 640x480, baseline 0.5..2 m,
                 rotation <= 20 deg).
  essential   -- as fundamental with f=1000 at 1024x768; returns K1, K2.

descriptor_ratio(seed, gt) draws the synthetic descriptor-ratio score of
each correspondence (lower is better, like a second-nearest-neighbour
ratio): inliers Beta(2, 5), outliers Beta(5, 2).  It uses its own stream
(derived from the scene seed), so adding it leaves every scene unchanged.
"""
import numpy as np

SEED_BASE = 0x67637273


def scene_seed(config_id, pair_idx):
    """Seed of pair ``pair_idx`` of workload ``config_id``."""
    return (SEED_BASE ^ (int(config_id) << 32) ^ int(pair_idx)) & ((1 << 63) - 1)


def _rot(axis, angle):
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K


def _finish(rng, inl, out, extra):
    corr = np.concatenate([inl, out], axis=0)
    gt = np.concatenate([np.ones(len(inl), bool), np.zeros(len(out), bool)])
    perm = rng.permutation(len(corr))
    d = dict(corr=np.ascontiguousarray(corr[perm]), gt=gt[perm])
    d.update(extra)
    return d


def homography_scene(seed, n=2000, inlier_ratio=0.5, noise=0.5, w=640, h=480):
    rng = np.random.default_rng(seed)
    th = np.deg2rad(rng.uniform(-30, 30))
    s = rng.uniform(0.7, 1.4)
    p = rng.normal(0.0, 1e-4, size=2)
    t = rng.uniform(-40, 40, size=2)
    A = np.array([[s * np.cos(th), -s * np.sin(th), t[0]],
                  [s * np.sin(th), s * np.cos(th), t[1]],
                  [p[0], p[1], 1.0]])
    Tc = np.array([[1, 0, w / 2], [0, 1, h / 2], [0, 0, 1.0]])
    H = Tc @ A @ np.linalg.inv(Tc)
    n_in = int(round(n * inlier_ratio))
    pts = []
    while len(pts) < n_in:
        x1 = rng.uniform([0, 0], [w, h], size=(n_in, 2))
        X = np.c_[x1, np.ones(n_in)] @ H.T
        x2 = X[:, :2] / X[:, 2:3] + rng.normal(0, noise, size=(n_in, 2))
        ok = (x2[:, 0] >= 0) & (x2[:, 0] < w) & (x2[:, 1] >= 0) & (x2[:, 1] < h) & (X[:, 2] > 0)
        pts.extend(np.c_[x1, x2][ok])
    inl = np.array(pts[:n_in]).reshape(-1, 4)
    out = np.c_[rng.uniform([0, 0], [w, h], size=(n - n_in, 2)), rng.uniform([0, 0], [w, h], size=(n - n_in, 2))]
    return _finish(rng, inl, out, dict(H=H, w1=w, h1=h, w2=w, h2=h))


def _two_view(rng, n, inlier_ratio, noise, f, w, h, plane_fraction=0.0):
    K = np.array([[f, 0, w / 2], [0, f, h / 2], [0, 0, 1.0]])
    R = _rot(rng.normal(size=3), np.deg2rad(rng.uniform(0, 20)))
    tdir = rng.normal(size=3)
    t = tdir / np.linalg.norm(tdir) * rng.uniform(0.5, 2.0)
    n_in = int(round(n * inlier_ratio))
    plane = None
    if plane_fraction > 0.0:
        # a dominant scene plane n . X = d (camera-1 frame) through depth ~6 m
        nrm = np.array([rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), 1.0])
        plane = (nrm / np.linalg.norm(nrm), rng.uniform(4.0, 8.0) / np.linalg.norm(nrm))
    pts = []
    while len(pts) < n_in:
        m = max(n_in, 256)
        u = rng.uniform([0, 0], [w, h], size=(m, 2))
        z = rng.uniform(2.0, 20.0, size=m)
        if plane is not None:
            ray = np.c_[(u[:, 0] - w / 2) / f, (u[:, 1] - h / 2) / f, np.ones(m)]
            on = rng.uniform(size=m) < plane_fraction
            z = np.where(on, plane[1] / (ray @ plane[0]), z)
        Xc = np.c_[(u[:, 0] - w / 2) / f * z, (u[:, 1] - h / 2) / f * z, z]
        X2 = Xc @ R.T + t
        ok = X2[:, 2] > 0.1
        u2 = X2[:, :2] / X2[:, 2:3] * f + np.array([w / 2, h / 2])
        u1n = u + rng.normal(0, noise, size=(m, 2))
        u2n = u2 + rng.normal(0, noise, size=(m, 2))
        ok &= (u2n[:, 0] >= 0) & (u2n[:, 0] < w) & (u2n[:, 1] >= 0) & (u2n[:, 1] < h)
        ok &= (u1n[:, 0] >= 0) & (u1n[:, 0] < w) & (u1n[:, 1] >= 0) & (u1n[:, 1] < h)
        pts.extend(np.c_[u1n, u2n][ok])
    inl = np.array(pts[:n_in]).reshape(-1, 4)
    out = np.c_[rng.uniform([0, 0], [w, h], size=(n - n_in, 2)), rng.uniform([0, 0], [w, h], size=(n - n_in, 2))]
    tx = np.array([[0, -t[2], t[1]], [t[2], 0, -t[0]], [-t[1], t[0], 0]])
    E = tx @ R
    Kinv = np.linalg.inv(K)
    F = Kinv.T @ E @ Kinv
    return inl, out, dict(E=E, F=F / np.linalg.norm(F), K1=K, K2=K.copy(), R=R, t=t, w1=w, h1=h, w2=w, h2=h)


def fundamental_scene(seed, n=5000, inlier_ratio=0.3, noise=0.5, f=800.0, w=640, h=480, plane_fraction=0.0):
    """plane_fraction > 0 puts that share of the inliers on one scene plane
    (a plane-dominated scene, where samples can be H-degenerate)."""
    rng = np.random.default_rng(seed)
    inl, out, extra = _two_view(rng, n, inlier_ratio, noise, f, w, h, plane_fraction)
    return _finish(rng, inl, out, extra)


def essential_scene(seed, n=10000, inlier_ratio=0.1, noise=0.75, f=1000.0, w=1024, h=768):
    rng = np.random.default_rng(seed)
    inl, out, extra = _two_view(rng, n, inlier_ratio, noise, f, w, h)
    return _finish(rng, inl, out, extra)


def descriptor_ratio(seed, gt):
    """Synthetic descriptor-ratio score: Beta(2, 5) for inliers, Beta(5, 2)
    for outliers (float64, in (0, 1))."""
    rng = np.random.default_rng([int(seed) & ((1 << 63) - 1), 0x726174696F])
    gt = np.asarray(gt, dtype=bool)
    r = np.empty(len(gt), np.float64)
    r[gt] = rng.beta(2.0, 5.0, size=int(gt.sum()))
    r[~gt] = rng.beta(5.0, 2.0, size=int((~gt).sum()))
    return r
