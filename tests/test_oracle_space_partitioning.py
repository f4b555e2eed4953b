"""The space partitioning inlier selector in the oracle
(inc/inlier_selectors/space_partitioning_ransac.h:339-504) and GC-RANSAC runs
that use it.  CPU only.

What pins it (closed forms the selector's rules give for known homographies):
- the identity on a grid with equal cell sizes in both images: a cell of 4 or
  more points is kept exactly when its second-image column is within one of
  its first-image column (the projected-corner test keeps the cells whose
  both second-image indices lie within one; the rectangle test that follows
  keeps a horizontal overlap whatever the rows, because its vertical clause
  can never hold); a mistake in either test, in which corners are tried, in
  the threshold growth or in the "fewer than 4 points" rule fails it;
- a translation by whole cells shifts that band by the translation;
- a singular homography (|det| < 1e-3) keeps nothing;
- the kept point count is the sum of the kept cells' sizes;
- the cell order is the grid container's iteration order (the oracle keeps
  the reference's container type and insertion order; the engine builds the
  same container).
And at run level: homography runs with the selector (with and without SPRT)
recover the plane.
"""
import numpy as np
import pytest

from pygcransac import synthetic


def _grid_points(rng, n, w, h):
    P = np.c_[rng.uniform(0, w, n), rng.uniform(0, h, n), rng.uniform(0, w, n), rng.uniform(0, h, n)]
    return P


@pytest.mark.parametrize("div", [4, 6, 8])
def test_identity_keeps_the_cells_within_one_column(oracle, div):
    rng = np.random.default_rng(div)
    w = h = 480.0
    P = _grid_points(rng, 6000, w, h)
    cs = np.full(4, w / div)
    out = oracle.sp_select(P, cs, div, np.eye(3), 4.5)
    ix1, iy1, ix2, iy2 = out["idx"].T
    expect = (out["npts"] < 4) | (np.abs(ix2 - ix1) <= 1)
    assert np.array_equal(out["kept"], expect)
    assert out["kept_points"] == out["npts"][out["kept"]].sum()
    # the corner test alone would also need the rows to agree: cells kept by
    # the rectangle test with rows far apart exist
    assert np.any(out["kept"] & (out["npts"] >= 4) & (np.abs(iy2 - iy1) > 1))


def test_translation_shifts_the_band(oracle):
    rng = np.random.default_rng(3)
    div, w = 6, 480.0
    cs = np.full(4, w / div)
    P = _grid_points(rng, 6000, w, w)
    T = np.array([[1, 0, 2 * cs[0]], [0, 1, 0], [0, 0, 1.0]])
    out = oracle.sp_select(P, cs, div, T, 4.5)
    ix1, iy1, ix2, iy2 = out["idx"].T
    expect = (out["npts"] < 4) | (np.abs(ix2 - (ix1 + 2)) <= 1)
    assert np.array_equal(out["kept"], expect)


def test_singular_homography_keeps_nothing(oracle):
    rng = np.random.default_rng(4)
    P = _grid_points(rng, 2000, 480.0, 480.0)
    cs = np.full(4, 120.0)
    S = np.array([[1, 2, 3], [2, 4, 6], [0, 0, 1.0]])
    out = oracle.sp_select(P, cs, 4, S, 4.5)
    assert out["kept_points"] == 0 and not out["kept"].any()


@pytest.mark.parametrize("use_sprt", [False, True])
def test_runs_with_space_partitioning_recover_the_plane(oracle, use_sprt):
    sc = synthetic.homography_scene(synthetic.scene_seed(16, 1), n=2500, inlier_ratio=0.5)
    p = oracle.params(3.0, 0.99, 0.975, 10000, 50, 0, 50, 4.0, 2, use_sprt=use_sprt, use_space_partitioning=True)
    rc, H, mask, st = oracle.find_homography(sc["corr"], 480, 640, 480, 640, p)
    assert rc > 0
    assert (mask & sc["gt"]).sum() >= 0.95 * sc["gt"].sum()
