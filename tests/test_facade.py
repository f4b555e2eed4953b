"""The C++ template API (include/gcransac.hpp) used as a reference program uses
GCRANSAC<Estimator, GridNeighborhoodGraph<4>> (src/pygcransac/include/GCRANSAC.h:
47-100; driver usage py/gcransac_python.cpp:1786-1830).

tests/cpp/facade_check.cpp drives the header exactly like such a program;
the GPU tests compare its runs with the oracle's GCRANSAC::run on the same
points, grid and seed: identical inlier sets and statistics, model within
1e-12 (the same bar as the find* drivers).  The CPU tests check that the
header compiles as C++17 and C++20 and that settings the engine does not
implement are rejected before any device work.
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from pygcransac import synthetic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "cpp", "libfacade_check.so")


class FacadeArgs(ctypes.Structure):
    _fields_ = [("threshold", ctypes.c_double), ("confidence", ctypes.c_double),
                ("spatial_coherence_weight", ctypes.c_double), ("max_iters", ctypes.c_int64),
                ("min_iters", ctypes.c_int64), ("sampler", ctypes.c_int32), ("lo_number", ctypes.c_int32),
                ("divisions", ctypes.c_int32), ("seed", ctypes.c_uint64), ("bad_setting", ctypes.c_int32),
                ("use_degensac", ctypes.c_int32), ("use_sprt", ctypes.c_int32), ("use_sp", ctypes.c_int32),
                ("sprt_ratio", ctypes.c_double), ("fps", ctypes.c_double)]


_lib = None


def _facade():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            pytest.fail("tests/cpp/libfacade_check.so is not built (run __graft_entry__.build())")
        _lib = ctypes.CDLL(LIB)
        _lib.facade_run.restype = ctypes.c_int
        _lib.facade_last_error.restype = ctypes.c_char_p
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def facade_run(kind, pts, grid_pts, cells, args, K=None):
    L = _facade()
    pts = np.ascontiguousarray(pts, np.float64)
    gp = None if grid_pts is None else np.ascontiguousarray(grid_pts, np.float64)
    cells = np.ascontiguousarray(cells, np.float64)
    Kc = None if K is None else np.ascontiguousarray(K, np.float64).ravel()
    n = pts.shape[0]
    M = np.zeros(9)
    mask = np.zeros(n, np.uint8)
    st = np.zeros(5, np.int64)
    score = ctypes.c_double(0.0)
    rc = L.facade_run(ctypes.c_int32(kind), _p(pts), ctypes.c_int64(n), _p(gp), _p(cells), ctypes.byref(args),
                      _p(Kc), _p(M), _p(mask), _p(st), ctypes.byref(score))
    return rc, M.reshape(3, 3), mask.astype(bool), st, score.value, L.facade_last_error().decode()


@pytest.mark.parametrize("std", ["c++17", "c++20"])
def test_header_compiles(std):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    subprocess.check_call([gxx, "-std=" + std, "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                           "-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "facade_check.cpp")])


@pytest.mark.parametrize("bad,needle", [(1, "max_graph_cut_number"), (3, "do_local_optimization")])
def test_unbuilt_settings_rejected(bad, needle):
    sc = synthetic.homography_scene(synthetic.scene_seed(0, 0), n=200)
    args = FacadeArgs(3.0, 0.99, 0.975, 1000, 50, 0, 50, 8, 0, bad, 1)
    rc, _, _, _, _, err = facade_run(0, sc["corr"], None, [80.0, 60.0, 80.0, 60.0], args)
    assert rc == -1
    assert needle in err


def _cells(w, h, div):
    return np.array([w / div, h / div, w / div, h / div])


def _compare(oracle, kind, pts, grid_pts, cells, thr, seed, sampler=0, K=None, degensac=True, sprt=0.0, sp=False,
             fps=0.0):
    args = FacadeArgs(thr, 0.99, 0.975, 100000, 50, sampler, 50, 8, seed, 0, 1 if degensac else 0,
                      1 if sprt > 0 else 0, 1 if sp else 0, sprt if sprt > 0 else 0.01, fps)
    p = oracle.params(thr, 0.99, 0.975, 100000, 50, sampler, 50, 8.0, seed, use_sprt=sprt > 0,
                      min_inlier_ratio_for_sprt=sprt if sprt > 0 else 1e-5, use_space_partitioning=sp)
    rco, Mo, mo, so = oracle.run(kind, pts, grid_pts, cells, p, 0 if degensac else 1)
    rc, M, mask, st, score, err = facade_run(kind, pts, grid_pts, cells, args, K)
    assert rc >= 0, err
    assert rc == rco
    assert np.array_equal(mask, mo)
    assert st[0] == so["iteration_number"]
    assert st[1] == so["local_optimization_number"]
    assert st[2] == so["graph_cut_number"]
    assert st[4] == (1 if sampler == 1 else 0)
    assert score == pytest.approx(so["score"], rel=1e-12, abs=1e-12)
    if rco > 0:
        np.testing.assert_allclose(M, Mo, rtol=1e-12, atol=1e-14)
    return rc


@pytest.mark.gpu
@pytest.mark.parametrize("seed,sampler", [(0, 0), (1, 0), (2, 1)])
def test_homography_matches_oracle(oracle, seed, sampler):
    sc = synthetic.homography_scene(synthetic.scene_seed(4, seed), n=1500, inlier_ratio=0.5)
    corr = sc["corr"]
    if sampler == 1:
        corr = corr[np.argsort(-sc["gt"].astype(int), kind="stable")]
    rc = _compare(oracle, 0, corr, None, _cells(640, 480, 8), 3.0, seed, sampler)
    assert rc > 0.4 * len(corr)


@pytest.mark.gpu
@pytest.mark.parametrize("sprt,sp", [(0.01, False), (0.0, True), (0.01, True)])
def test_homography_preemption_and_selector_slots_match_oracle(oracle, sprt, sp):
    # GCRANSAC<..., SPRTPreemptiveVerfication, SpacePartitioningRANSAC>
    # (GCRANSAC.h:50-51) on the caller's grid
    sc = synthetic.homography_scene(synthetic.scene_seed(4, 30), n=2000, inlier_ratio=0.45)
    rc = _compare(oracle, 0, sc["corr"], None, _cells(640, 480, 8), 3.0, 5, sprt=sprt, sp=sp)
    assert rc > 0.35 * len(sc["corr"])


@pytest.mark.gpu
@pytest.mark.parametrize("kind,ratio", [(1, 0.05), (2, 0.1)])
def test_sprt_slot_matches_oracle_f_e(oracle, kind, ratio):
    if kind == 1:
        sc = synthetic.fundamental_scene(synthetic.scene_seed(4, 31), n=2000, inlier_ratio=0.5)
        rc = _compare(oracle, 1, sc["corr"], None, _cells(640, 480, 8), 1.0, 6, sprt=ratio)
    else:
        sc = synthetic.essential_scene(synthetic.scene_seed(4, 32), n=2000, inlier_ratio=0.5)
        pts = oracle.normalise_correspondences(sc["corr"], sc["K1"], sc["K2"])
        K1, K2 = sc["K1"], sc["K2"]
        normaliser = 0.25 * (K1[0, 0] + K1[1, 1] + K2[0, 0] + K2[1, 1])
        K = np.concatenate([np.ravel(K1), np.ravel(K2)])
        rc = _compare(oracle, 2, pts, sc["corr"], _cells(1024, 768, 8), 0.75 / normaliser, 6, K=K, sprt=ratio)
    assert rc > 0.35 * 2000


@pytest.mark.gpu
def test_setfps_generous_limit_matches_oracle(oracle):
    # GCRANSAC::setFPS (GCRANSAC.h:86, 503-516): a limit the run never
    # reaches leaves it unchanged
    sc = synthetic.homography_scene(synthetic.scene_seed(4, 34), n=1500, inlier_ratio=0.5)
    _compare(oracle, 0, sc["corr"], None, _cells(640, 480, 8), 3.0, 7, fps=1e-4)


@pytest.mark.gpu
def test_setfps_tiny_limit_stops_after_the_first_iteration():
    sc = synthetic.fundamental_scene(synthetic.scene_seed(4, 35), n=2000, inlier_ratio=0.3)
    args = FacadeArgs(1.0, 0.99, 0.975, 100000, 50, 0, 50, 8, 3, 0, 1, 0, 0, 0.01, 1e9)
    rc, _, _, st, _, err = facade_run(1, sc["corr"], None, _cells(640, 480, 8), args)
    assert rc >= 0, err
    # the first iteration (and the draws it needed) then the time check
    assert 1 <= st[0] <= 100
    args.fps = 0.0
    rc2, _, _, st2, _, _ = facade_run(1, sc["corr"], None, _cells(640, 480, 8), args)
    assert st2[0] > 50


@pytest.mark.gpu
@pytest.mark.parametrize("kind", [0, 1])
def test_user_sampler_plugin_matches_builtin_uniform(kind):
    # A sampler class of the caller's own (samplers/sampler.h:44-87) that
    # restates the reference's uniform sampler with std::mt19937 and
    # std::uniform_int_distribution draws the main loop's samples through the
    # engine's callback; seeded like the built-in uniform sampler it must give
    # the built-in run bit for bit (a dropped, repeated or reordered callback
    # draw changes the samples and so the run).
    if kind == 0:
        sc = synthetic.homography_scene(synthetic.scene_seed(4, 40), n=1500, inlier_ratio=0.4)
        thr = 3.0
    else:
        sc = synthetic.fundamental_scene(synthetic.scene_seed(4, 41), n=2000, inlier_ratio=0.5)
        thr = 1.0
    out = []
    for sampler in (0, 2):
        args = FacadeArgs(thr, 0.99, 0.975, 100000, 50, sampler, 50, 8, 11, 0, 1, 0, 0, 0.01, 0.0)
        out.append(facade_run(kind, sc["corr"], None, _cells(640, 480, 8), args))
    (rc0, M0, m0, st0, s0, e0), (rc1, M1, m1, st1, s1, e1) = out
    assert rc0 > 0.3 * len(sc["corr"]), e0
    assert rc1 == rc0, e1
    assert np.array_equal(m1, m0)
    assert np.array_equal(st1[:4], st0[:4])
    assert st0[4] == 0 and st1[4] == 2
    assert s1 == s0
    assert np.array_equal(M1.view(np.int64), M0.view(np.int64))


@pytest.mark.gpu
def test_user_sampler_exception_reaches_the_caller():
    sc = synthetic.homography_scene(synthetic.scene_seed(4, 42), n=800, inlier_ratio=0.4)
    args = FacadeArgs(3.0, 0.99, 0.975, 100000, 50, 3, 50, 8, 11, 0, 1, 0, 0, 0.01, 0.0)
    rc, _, _, _, _, err = facade_run(0, sc["corr"], None, _cells(640, 480, 8), args)
    assert rc == -2 and "user sampler failure" in err


def test_space_partitioning_slot_rejected_for_fundamental():
    sc = synthetic.fundamental_scene(synthetic.scene_seed(4, 33), n=300)
    args = FacadeArgs(1.0, 0.99, 0.975, 1000, 50, 0, 50, 8, 0, 0, 1, 0, 1, 0.01, 0.0)
    rc, _, _, _, _, err = facade_run(1, sc["corr"], None, _cells(640, 480, 8), args)
    assert rc == -1 and "homographies" in err


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [0, 1])
def test_fundamental_matches_oracle(oracle, seed):
    sc = synthetic.fundamental_scene(synthetic.scene_seed(4, 10 + seed), n=2000, inlier_ratio=0.5)
    rc = _compare(oracle, 1, sc["corr"], None, _cells(640, 480, 8), 1.0, seed)
    assert rc > 0.4 * len(sc["corr"])


@pytest.mark.gpu
@pytest.mark.parametrize("seed,degensac", [(0, True), (0, False), (2, True)])
def test_fundamental_plane_scene_matches_oracle(oracle, seed, degensac):
    # plane-dominated: DEGENSAC fires when the estimator has it switched on
    sc = synthetic.fundamental_scene(synthetic.scene_seed(7, seed), n=1500, inlier_ratio=0.5, plane_fraction=0.8)
    rc = _compare(oracle, 1, sc["corr"], None, _cells(640, 480, 8), 1.0, seed, degensac=degensac)
    assert rc > 0.3 * len(sc["corr"])


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [0, 1])
def test_essential_matches_oracle(oracle, seed):
    sc = synthetic.essential_scene(synthetic.scene_seed(4, 20 + seed), n=2000, inlier_ratio=0.5)
    pts = oracle.normalise_correspondences(sc["corr"], sc["K1"], sc["K2"])
    K1, K2 = sc["K1"], sc["K2"]
    normaliser = 0.25 * (K1[0, 0] + K1[1, 1] + K2[0, 0] + K2[1, 1])
    K = np.concatenate([np.ravel(K1), np.ravel(K2)])
    rc = _compare(oracle, 2, pts, sc["corr"], _cells(1024, 768, 8), 0.75 / normaliser, seed, K=K)
    assert rc > 0.4 * len(pts)


@pytest.mark.gpu
def test_no_model_leaves_empty_descriptor(oracle):
    rng = np.random.default_rng(5)
    corr = np.column_stack([rng.uniform(0, 640, 300), rng.uniform(0, 480, 300),
                            rng.uniform(0, 640, 300), rng.uniform(0, 480, 300)])
    args = FacadeArgs(0.01, 0.99, 0.975, 200, 50, 0, 50, 8, 3, 0, 1)
    p = oracle.params(0.01, 0.99, 0.975, 200, 50, 0, 50, 8.0, 3)
    rco, _, mo, so = oracle.run(0, corr, None, _cells(640, 480, 8), p)
    rc, M, mask, st, _, err = facade_run(0, corr, None, _cells(640, 480, 8), args)
    assert rc == rco == 0, err
    assert not mask.any() and not mo.any()
    assert st[0] == so["iteration_number"]


def test_host_random_stream_matches_std(tmp_path):
    # the engine's sampler stream defines parity with the reference's runs:
    # its MT19937 must equal std::mt19937 and the unique-index draws must be
    # the ones std::uniform_int_distribution makes from it
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path / "rng_check")
    subprocess.check_call([gxx, "-std=c++17", "-O2", "-o", exe, os.path.join(ROOT, "tests", "cpp", "rng_check.cpp")])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip() == "ok"


def test_host_pool_runs_every_item_once(tmp_path):
    # the engine's host thread pool: item coverage, idle gaps, two pools at
    # once, exceptions
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path / "pool_check")
    subprocess.check_call([gxx, "-std=c++20", "-O2", "-pthread", "-o", exe,
                           os.path.join(ROOT, "tests", "cpp", "pool_check.cpp")])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120, env=dict(os.environ, GCR_HOST_THREADS="8"))
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip() == "ok"
