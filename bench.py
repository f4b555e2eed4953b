#!/usr/bin/env python3
"""Benchmark: image pairs/s of Graph-Cut RANSAC on one MI355X.

Default workload (the headline metric's configuration): learned-guidance
GC-RANSAC on 1,024 held-out fundamental-matrix pairs per step -- f=800
pinhole cameras, 640x480, 4,000 correspondences at 5 % inliers with 0.5 px
noise, synthetic descriptor-ratio scores, 1 px Sampson threshold, grid
neighbourhood (8 divisions), spatial coherence weight 0.975, confidence
0.99, max_iters 10^4 (the API default; at 5 % inliers the 0.99-confidence
rule would ask for ~6e9 seven-point samples).  A step is one
gcr_find_batch_guided call over the batch (features, network, per-pair sort
and PROSAC GC-RANSAC), with correspondences and scores resident in HBM before
the timed region.

Reported next to the throughput:
  * the plain GPU path (uniform sampler, one pair per engine run) over the
    first --plain-pairs pairs of the same batch: pairs/s, recall of the
    generator's inliers and inlier counts, beside the guided run's;
  * equal_inlier_comparison: a workload where the plain path does reach the
    0.99-confidence rule (fundamental, 5,000 correspondences at 30 %
    inliers, max_iters 10^6): guided+batched against plain pairs/s, and how
    far the final inlier counts of the two paths lie apart, pair by pair;
  * fundamental_batched: the same fundamental workload batched with the
    uniform sampler (no guidance);
  * roofline of the dominant kernel family by HIP-event time on the engine
    streams over the timed steps (achieved = its algorithmic work over its
    kernel time), the scoring kernel's roofline, hyp*pts/s of scoring;
  * cpu_baseline: the C++ oracle restatement of the reference (uniform
    sampler, the reference has no learned guidance) on the same workload,
    median of 5 repetitions of a fixed sample of pairs, on all the host
    threads this process may use (one pair per thread) and on one thread.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--pairs B]
                       [--workload guided|fundamental|homography|mixed|essential] [--no-cpu]

Other workloads (--workload): fundamental (5,000 correspondences, 30 %
inliers, max_iters 10^6), homography (2,000 correspondences, 50 %, 3 px,
max_iters 10^6), mixed (1,024 pairs: half homography at 3 px, half
fundamental at 1 px, 200-8,000 correspondences log-uniform, inlier ratios
uniform in [0.1, 0.7], max_iters 10^4), essential (f=1000 at 1024x768,
10,000 correspondences at 10 % inliers, 0.75 px, max_iters 10^6: runs to the
0.99-confidence rule).
"""
import argparse
import ctypes
import json
import os
import statistics
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "graph-cut-ransac_amd"))
# The batched engine runs its concurrent groups on separate HIP streams; HIP
# maps a process's streams onto GPU_MAX_HW_QUEUES hardware queues and
# streams sharing a queue serialise.  This benchmark runs with 8 queues (the
# engine's recommended setting, also what pygcransac asks for when the
# variable is unset) whatever the environment's value, and reports it.
_HW_QUEUES_ENV = os.environ.get("GPU_MAX_HW_QUEUES")
os.environ["GPU_MAX_HW_QUEUES"] = "8"

import numpy as np  # noqa: E402

from pygcransac import _native, hipmem, synthetic  # noqa: E402

PEAK = {"fp64-valu": (78.6, "TFLOP/s"), "bf16-mfma": (2500.0, "TFLOP/s"), "fp16-mfma": (2500.0, "TFLOP/s"),
        "hbm": (8000.0, "GB/s")}

FIT_KERNEL = {0: "k_fit_h", 1: "k_fit_f", 2: "k_fit_e"}

# HBM traffic per launch of the fit kernels (KiB read, KiB written), from
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of this script for each
# workload (one timed step, the batch groups pinned by GCR_BATCH_GROUPS),
# scaled by the counters' reading for a 256 MiB device copy (FETCH_SIZE
# x2.0, WRITE_SIZE x1.0), averaged over the launches of k_fit_f<64> and
# k_fit_f<256> (k_fit_h<64> and <256> for homographies).  Key: (workload,
# pairs per step, batch groups).  The summary with the commands is
# profiles/pmc_traffic_fit_kernels.md.
TRAFFIC_KIB = {
    ("guided", 1024, 2): (2479.6, 44.6),
    ("guided", 1024, 3): (2022.3, 34.1),
    ("fundamental", 1024, 2): (8523.2, 149.6),
    ("fundamental", 1024, 3): (6292.7, 107.6),
    ("guided", 1024, 4): (1729.1, 28.3),
    ("fundamental", 1024, 4): (5261.4, 88.5),
    ("homography", 1024, 2): (8342.1, 171.2),
    ("homography", 1024, 4): (4776.2, 95.8),
}
# Algorithmic HBM bytes of a non-minimal fit per point: the gathered
# correspondence (4 doubles) and its index.
FIT_BYTES_PER_POINT = 36.0

WORKLOADS = {
    # name: (kind, generator, n, inlier_ratio, noise, threshold, image (w, h), config id)
    "fundamental": (1, synthetic.fundamental_scene, 5000, 0.3, 0.5, 1.0, (640, 480), 1),
    "homography": (0, synthetic.homography_scene, 2000, 0.5, 0.5, 3.0, (640, 480), 0),
    # held-out pairs of the learned-guidance evaluation: 4,000 correspondences at 5 % inliers
    "guided": (1, synthetic.fundamental_scene, 4000, 0.05, 0.5, 1.0, (640, 480), 4),
    # mixed batch: half homography pairs (threshold 3 px), half fundamental
    # pairs (1 px), correspondence counts log-uniform in [200, 8000], inlier
    # ratios uniform in [0.1, 0.7]; kind, n and ratio are drawn per pair
    "mixed": (-1, None, 0, 0.0, 0.5, None, (640, 480), 3),
    "essential": (2, synthetic.essential_scene, 10000, 0.1, 0.75, 0.75, (1024, 768), 2),
}

# max_iters per workload: fundamental, homography and essential pairs run to
# the 0.99-confidence rule; at 5 % inliers the rule asks for ~6e9 seven-point
# samples, so the guided workload runs at the API default cap of 10^4, and so
# does the mixed batch (its fundamental pairs at 10 % inliers would ask for
# ~5e7); --max-iters overrides
MAX_ITERS = {"fundamental": 1000000, "homography": 1000000, "guided": 10000, "mixed": 10000,
             "essential": 1000000}
DEFAULT_PAIRS = {"essential": 32}
THRESHOLD = {0: 3.0, 1: 1.0, 2: 0.75}
FLOP_PER_HYP_PT = {0: 18.0, 1: 42.0, 2: 42.0}
SCORE_FAMILY = {0: "score_h", 1: "score_f", 2: "score_e"}
SOLVER_FAMILY = {"solve_h4": "k_h4_solve (four-point homography)", "solve_f7": "k_f7_solve (seven-point fundamental)",
                 "solve_e5": "k_e5_solve (five-point essential, Nister + Sturm)"}


def make_pairs(name, count, first=0):
    kind, gen, n, eps, noise, thr, (w, h), cid = WORKLOADS[name]
    scenes = []
    for i in range(count):
        seed = synthetic.scene_seed(cid, first + i)
        k, g, ni, ei = kind, gen, n, eps
        if name == "mixed":
            rng = np.random.default_rng(seed ^ 0x6D6978)
            k = (first + i) % 2
            g = synthetic.fundamental_scene if k == 1 else synthetic.homography_scene
            ni = int(np.floor(np.exp(rng.uniform(np.log(200.0), np.log(8000.0)))))
            ei = float(rng.uniform(0.1, 0.7))
        sc = g(seed, n=ni, inlier_ratio=ei, noise=noise)
        sc.update(seed=seed, kind=k, thr=THRESHOLD[k])
        scenes.append(sc)
    return scenes


def params_for(max_iters, seed, kind, thr, sampler=0):
    p = _native.GcrParams()
    _native.lib().gcr_default_params(ctypes.byref(p), kind)
    p.threshold = thr
    p.confidence = 0.99
    p.spatial_coherence_weight = 0.975
    p.max_iters = max_iters
    p.min_iters = 50
    p.sampler = sampler
    p.use_sprt = 0
    p.use_space_partitioning = 0
    p.neighborhood = 0
    p.neighborhood_size = 8.0
    p.lo_number = 50
    p.seed = seed
    return p


class Batch:
    """gcr_find_batch over a fixed set of pairs, correspondences in HBM."""

    def __init__(self, name, scenes, max_iters, sampler=0):
        _, _, _, _, _, _, (w, h), _ = WORKLOADS[name]
        self.n_pairs = len(scenes)
        corr = np.ascontiguousarray(np.concatenate([s["corr"] for s in scenes]))
        self.offsets = np.zeros(self.n_pairs + 1, np.int64)
        self.offsets[1:] = np.cumsum([len(s["corr"]) for s in scenes])
        self.kinds = np.array([s["kind"] for s in scenes], np.int32)
        self.sizes = np.tile(np.array([h, w, h, w], np.int32), self.n_pairs)
        self.K = None
        if any(s["kind"] == 2 for s in scenes):
            self.K = np.zeros(18 * self.n_pairs)
            for i, s in enumerate(scenes):
                if s["kind"] == 2:
                    self.K[18 * i:18 * i + 9] = s["K1"].ravel()
                    self.K[18 * i + 9:18 * i + 18] = s["K2"].ravel()
        self.params = (_native.GcrParams * self.n_pairs)()
        for i in range(self.n_pairs):
            self.params[i] = params_for(max_iters, i, scenes[i]["kind"], scenes[i]["thr"], sampler)
        self.dev = hipmem.DeviceArray(corr)
        self.models = np.zeros(9 * self.n_pairs)
        self.masks = np.zeros(int(self.offsets[-1]), np.uint8)
        self.ninl = np.zeros(self.n_pairs, np.int32)
        self.stats = (_native.GcrStats * self.n_pairs)()

    def run(self, lo=0, hi=None):
        hi = self.n_pairs if hi is None else hi
        L = _native.lib()
        vp = ctypes.c_void_p
        off = self.offsets[lo:hi + 1].copy()
        rc = L.gcr_find_batch(
            hi - lo, self.kinds[lo:hi].ctypes.data_as(vp), off.ctypes.data_as(vp),
            ctypes.c_void_p(self.dev.ptr.value), 1, self.sizes[4 * lo:4 * hi].ctypes.data_as(vp),
            None if self.K is None else self.K[18 * lo:].ctypes.data_as(vp),
            ctypes.cast(ctypes.byref(self.params, lo * ctypes.sizeof(_native.GcrParams)), vp),
            self.models[9 * lo:].ctypes.data_as(vp),
            self.masks[int(self.offsets[lo]):].ctypes.data_as(vp), self.ninl[lo:].ctypes.data_as(vp),
            ctypes.cast(ctypes.byref(self.stats, lo * ctypes.sizeof(_native.GcrStats)), vp))
        _native.raise_for(rc)

    def inliers(self):
        return np.array([self.stats[i].num_inliers for i in range(self.n_pairs)])

    def free(self):
        self.dev.free()


class GuidedBatch:
    """The learned-guidance path over the same pairs: features, network,
    per-pair sort and PROSAC GC-RANSAC in one gcr_find_batch_guided call,
    with correspondences and descriptor-ratio scores resident in HBM."""

    def __init__(self, name, scenes, net, max_iters):
        self.plain = Batch(name, scenes, max_iters, sampler=1)
        ratio = np.ascontiguousarray(np.concatenate([synthetic.descriptor_ratio(s["seed"], s["gt"])
                                                     for s in scenes]))
        self.dev_ratio = hipmem.DeviceArray(ratio)
        self.net = np.ascontiguousarray(net, np.float32)
        self.n_pairs = self.plain.n_pairs

    def run(self):
        b = self.plain
        vp = ctypes.c_void_p
        rc = _native.lib().gcr_find_batch_guided(
            b.n_pairs, b.kinds.ctypes.data_as(vp), b.offsets.ctypes.data_as(vp), ctypes.c_void_p(b.dev.ptr.value),
            1, b.sizes.ctypes.data_as(vp), None, ctypes.c_void_p(self.dev_ratio.ptr.value),
            self.net.ctypes.data_as(vp), ctypes.cast(b.params, vp), b.models.ctypes.data_as(vp),
            b.masks.ctypes.data_as(vp), b.ninl.ctypes.data_as(vp), ctypes.cast(b.stats, vp), None)
        _native.raise_for(rc)

    @property
    def masks(self):
        return self.plain.masks

    @property
    def offsets(self):
        return self.plain.offsets

    def inliers(self):
        return self.plain.inliers()

    def free(self):
        self.dev_ratio.free()
        self.plain.free()


def recall(batch, scenes, count=None):
    """Per pair: fraction of the generator's true inliers in the returned mask."""
    out = []
    for i, s in enumerate(scenes[:count]):
        m = batch.masks[int(batch.offsets[i]):int(batch.offsets[i + 1])].astype(bool)
        out.append((m & s["gt"]).sum() / max(1, s["gt"].sum()))
    return np.array(out)


def timed(fn, steps):
    hipmem.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    hipmem.synchronize()
    return time.perf_counter() - t0


def plain_path(batch, count):
    """One pair per engine run (uniform sampler) over the first `count` pairs."""
    t = timed(lambda: [batch.run(i, i + 1) for i in range(count)], 1)
    return count / t


def host_info():
    info = {"os_cpu_count": os.cpu_count()}
    try:
        info["affinity_threads"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    try:
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=10).stdout.strip())
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info["lscpu " + k.strip()] = v.strip()
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline(name, scenes, max_iters, reps=5):
    """The C++ oracle restatement of the reference (uniform sampler) over a
    fixed sample of `scenes`: the median of `reps` repetitions, once with one
    pair per thread on every host thread this process may use (at most 16,
    the box's share for one GPU) and once on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py  # the C++ oracle restatement (test infrastructure)

    oracle_py.lib()
    _, _, _, _, _, _, (w, h), _ = WORKLOADS[name]

    def one(i):
        sc = scenes[i]
        p = oracle_py.params(sc["thr"], 0.99, 0.975, max_iters, 50, 0, 50, 8.0, i)
        if sc["kind"] == 1:
            oracle_py.find_fundamental(sc["corr"], h, w, h, w, p)
        elif sc["kind"] == 2:
            oracle_py.find_essential(sc["corr"], sc["K1"], sc["K2"], h, w, h, w, p)
        else:
            oracle_py.find_homography(sc["corr"], h, w, h, w, p)

    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        threads = os.cpu_count() or 1
    threads = max(1, min(16, threads))
    # size the sample so one single-thread repetition takes about 2 s; a
    # workload whose single pair takes longer than 5 s (essential pairs of
    # 10,000 points at 10 % inliers) is timed once per mode
    t0 = time.perf_counter()
    one(0)
    per_pair = max(1e-3, time.perf_counter() - t0)
    print("# cpu baseline: one pair in %.2f s" % per_pair, file=sys.stderr, flush=True)
    if per_pair > 5.0:
        reps, n1, nall = 1, 1, min(len(scenes), threads)
    else:
        n1 = int(max(2, min(len(scenes), 2.0 / per_pair)))
        nall = int(min(len(scenes), max(threads, n1 * threads)))
    single, multi = [], []
    for r in range(reps):
        t0 = time.perf_counter()
        for i in range(n1):
            one(i)
        single.append(n1 / (time.perf_counter() - t0))
        print("# cpu baseline: single-thread repetition %d done" % r, file=sys.stderr, flush=True)
    with ThreadPoolExecutor(max_workers=threads) as ex:
        for r in range(reps):
            t0 = time.perf_counter()
            list(ex.map(one, range(nall)))
            multi.append(nall / (time.perf_counter() - t0))
            print("# cpu baseline: %d-thread repetition %d done" % (threads, r), file=sys.stderr, flush=True)
    return {"value": statistics.median(multi), "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": "oracle restatement of the reference CPU path (uniform sampler, max_iters %d): median of %d "
                      "repetitions over the first %d pairs of a held-out %s batch, one pair per thread on %d "
                      "threads (ctypes releases the GIL)" % (max_iters, reps, nall, name, threads),
            "single_thread": {"value": statistics.median(single), "unit": "pairs/s", "cores": 1,
                              "sample": "median of %d repetitions over the first %d pairs" % (reps, n1)},
            "host": host_info()}


def profile_summary(steps, kinds_run):
    host = {}
    for ph in ("setup", "start", "copies", "chunk", "label", "fit", "masklist", "sprt", "ahead", "finish", "resume",
               "syncwait"):
        r = _native.profile_read("host_" + ph)
        if r:
            host[ph] = round(r[0] / steps, 3)
    kern, fams = {}, {}
    for fam_name in ("solve_h4", "solve_f7", "solve_e5", "score_h", "score_f", "score_e", "energy", "mincut", "fit"):
        r = _native.profile_read(fam_name)
        if r:
            kern[fam_name] = round(r[0] / steps, 3)
            fams[fam_name] = r
    fit_parts = {}
    for sub in ("fit_lo", "fit_irls", "fit_post"):
        r = _native.profile_read(sub)
        if r:
            fit_parts[sub] = {"ms_per_step": round(r[0] / steps, 3), "launches_per_step": r[1] / steps,
                              "avg_launch_ms": r[0] / max(1, r[1]), "points_per_launch": r[2] / max(1, r[1])}
    fam_of = {k: SCORE_FAMILY[k] for k in kinds_run}
    sc_ms = sc_launch = sc_hp = sc_flop = 0.0
    for fam in sorted(set(fam_of.values())):
        r = fams.get(fam)
        if not r:
            continue
        sc_ms += r[0]
        sc_launch += r[1]
        sc_hp += r[2]
        sc_flop += FLOP_PER_HYP_PT[0 if fam == "score_h" else 1] * r[2]
    return host, kern, fams, fit_parts, sc_ms, sc_launch, sc_hp, sc_flop


def roofline_entry(kernel, bound, achieved, launches, avg_ms, **extra):
    peak, unit = PEAK[bound]
    d = {"kernel": kernel, "bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
         "frac": achieved / peak, "avg_launch_ms": avg_ms, "launches": launches}
    d.update(extra)
    return d


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=4,
                    help="untimed steps (the first four calls of a batch shape also settle its group count)")
    ap.add_argument("--pairs", type=int, default=None, help="pairs per step (default 1,024; essential 32)")
    ap.add_argument("--workload", default="guided", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the plain-path and equal-inlier comparisons")
    ap.add_argument("--max-iters", type=int, default=None, help="override the workload's max_iters")
    ap.add_argument("--net", default=None, help="network parameter file of the guided workload "
                    "(default: the committed pygcransac/data/guided_mlp.npz)")
    ap.add_argument("--plain-pairs", type=int, default=32)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run one device copy of this many MiB (profiler counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box in this setup has one MI355X; "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration_mib > 0:
        nbytes = args.copy_calibration_mib << 20
        src, dst = hipmem.DeviceBuffer(nbytes), hipmem.DeviceBuffer(nbytes)
        hipmem.copy_d2d(dst, src, nbytes)
        hipmem.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes}))
        return
    L = _native.lib()
    if L.gcr_device_count() <= 0:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    name = args.workload
    kind, _, n, eps, noise, thr, (w, h), cid = WORKLOADS[name]
    max_iters = args.max_iters or MAX_ITERS[name]
    pairs = args.pairs or DEFAULT_PAIRS.get(name, 1024)
    scenes = make_pairs(name, pairs)
    if name == "mixed":
        n = sum(len(s["corr"]) for s in scenes)
    net = None
    if name == "guided":
        from pygcransac import guided
        net = guided.load(args.net) if args.net else guided.load()
        runner = GuidedBatch(name, scenes, net, max_iters)
    else:
        runner = Batch(name, scenes, max_iters)
    print("# workload %s: %d pairs, %d correspondences %s per step, max_iters %d" %
          (name, pairs, n, "in all" if name == "mixed" else "per pair", max_iters), file=sys.stderr, flush=True)
    for i in range(args.warmup):
        runner.run()
        print("# warmup step %d done" % i, file=sys.stderr, flush=True)
    hipmem.synchronize()
    _native.profile_reset()
    _native.profile_enable(True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        runner.run()
        print("# step %d done" % i, file=sys.stderr, flush=True)
    hipmem.synchronize()
    dt = time.perf_counter() - t0
    _native.profile_enable(False)
    # concurrent batch groups of the last timed step (the library's batch-shape
    # autotuner settles on a count in the first four calls of a shape)
    groups = int(_native.lib().gcr_last_batch_groups())
    kinds_run = sorted(set(int(s["kind"]) for s in scenes))
    host, kern, fams, fit_parts, sc_ms, sc_launch, sc_hp, sc_flop = profile_summary(args.steps, kinds_run)
    fit_ms, fit_launch, fit_flops = _native.profile_read("fit") or (0.0, 0, 0.0)
    dominant = max(fams, key=lambda k: fams[k][0]) if fams else "fit"
    value = pairs * args.steps / dt
    run_inl = runner.inliers()
    run_recall = recall(runner, scenes)
    iters = np.array([(runner.plain if name == "guided" else runner).stats[i].iteration_number
                      for i in range(pairs)])
    sampler_txt = ("PROSAC over the learned network's order (features, network and sort inside the timed region)"
                   if name == "guided" else "uniform sampler")
    if name == "mixed":
        workload = ("mixed GC-RANSAC batch, %d pairs per step (half homography at 3 px, half fundamental at 1 px), "
                    "%d correspondences in all (per pair log-uniform in [200, 8000]), inlier ratios uniform in "
                    "[0.1, 0.7], noise 0.5 px, grid 8 divisions, lambda 0.975, conf 0.99, max_iters %d, uniform "
                    "sampler" % (pairs, n, max_iters))
    else:
        workload = ("%s GC-RANSAC, %d pairs per step, N=%d, inlier ratio %.2f, noise %.2f px, threshold %.2f px, "
                    "%dx%d images, grid 8 divisions, lambda 0.975, conf 0.99, max_iters %d, %s" %
                    ("fundamental" if name == "guided" else name, pairs, n, eps, noise, thr, w, h, max_iters,
                     sampler_txt))
    out = {
        "metric": "image pairs/s to 0.99 conf, guided+batched vs plain GPU; hyp*pts/s scoring",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload, "pairs_per_step": pairs, "correspondences": n, "max_iters": max_iters,
                   "batch_groups": groups, "gpu_max_hw_queues": int(os.environ["GPU_MAX_HW_QUEUES"]),
                   "gpu_max_hw_queues_in_environment": _HW_QUEUES_ENV},
        "mean_true_inlier_recall": float(run_recall.mean()),
        "mean_inliers": float(run_inl.mean()),
        "mean_iterations": float(iters.mean()),
    }
    # ---- plain GPU path on the same pairs (uniform sampler, one pair per call)
    if args.plain_pairs > 0 and not args.no_extras:
        cnt = min(args.plain_pairs, pairs)
        plain_batch = runner.plain if name == "guided" else runner
        if name == "guided":
            plain_batch = Batch(name, scenes[:cnt], max_iters, sampler=0)
        guided_rec = run_recall[:cnt]
        guided_inl = run_inl[:cnt]
        pps = plain_path(plain_batch, cnt)
        prec = recall(plain_batch, scenes, cnt)
        pinl = plain_batch.inliers()[:cnt]
        out["plain_gpu"] = {
            "pairs_per_s": pps, "pairs": cnt,
            "path": "one pair per engine run, uniform sampler, same pairs, same max_iters",
            "mean_true_inlier_recall": float(prec.mean()), "mean_inliers": float(pinl.mean()),
            "this_run_on_same_pairs": {"mean_true_inlier_recall": float(guided_rec.mean()),
                                       "mean_inliers": float(guided_inl.mean())},
            "speedup": value / pps,
        }
        if name == "guided":
            plain_batch.free()
    # ---- equal final inliers where the plain path reaches 0.99 confidence
    if name == "guided" and not args.no_extras:
        eq_scenes = make_pairs("fundamental", 256, first=200000)
        eq_iters = MAX_ITERS["fundamental"]
        gb = GuidedBatch("fundamental", eq_scenes, net, eq_iters)
        gb.run()
        t_g = timed(gb.run, 1)
        g_inl = gb.inliers()
        g_rec = recall(gb, eq_scenes)
        g_it = np.array([gb.plain.stats[i].iteration_number for i in range(len(eq_scenes))])
        cnt = min(args.plain_pairs, 32)
        pb = Batch("fundamental", eq_scenes[:cnt], eq_iters, sampler=0)
        p_pps = plain_path(pb, cnt)
        p_inl = pb.inliers()[:cnt]
        p_it = np.array([pb.stats[i].iteration_number for i in range(cnt)])
        rel = np.abs(g_inl[:cnt] - p_inl) / np.maximum(1, p_inl)
        ub = Batch("fundamental", eq_scenes, eq_iters, sampler=0)
        ub.run()
        t_u = timed(ub.run, 1)
        out["equal_inlier_comparison"] = {
            "workload": "fundamental GC-RANSAC, %d held-out pairs, N=5000, inlier ratio 0.30, noise 0.5 px, "
                        "threshold 1 px, max_iters %d (every plain run stops at the 0.99-confidence rule)"
                        % (len(eq_scenes), eq_iters),
            "guided_batched_pairs_per_s": len(eq_scenes) / t_g,
            "plain_gpu_pairs_per_s": p_pps,
            "speedup": len(eq_scenes) / t_g / p_pps,
            "pairs_compared": cnt,
            "pairs_within_0.5pct_inliers": int((rel <= 0.005).sum()),
            "mean_abs_rel_inlier_difference": float(rel.mean()),
            "guided_mean_inliers": float(g_inl[:cnt].mean()), "plain_mean_inliers": float(p_inl.mean()),
            "guided_mean_recall": float(g_rec.mean()),
            "guided_mean_iterations": float(g_it.mean()), "plain_mean_iterations": float(p_it.mean()),
            "plain_runs_below_max_iters": int((p_it < eq_iters).sum()),
            "uniform_batched_pairs_per_s": len(eq_scenes) / t_u,
        }
        gb.free()
        pb.free()
        ub.free()
    # ---- rooflines
    out["host_ms_per_step"] = host
    out["kernel_ms_per_step"] = kern
    out["fit_kernel_breakdown"] = fit_parts
    out["scoring_hyp_pts_per_s"] = sc_hp / (sc_ms * 1e-3) if sc_ms > 0 else 0.0
    score_ach = sc_flop / (sc_ms * 1e-3) / 1e12 if sc_ms > 0 else 0.0
    # candidate chunks score on the matrix cores (rejection) plus the FP64
    # exact pass unless GCR_SCORE_MFMA=0 selects the FP64 kernel
    mx_on = os.environ.get("GCR_SCORE_MFMA", "1") != "0"
    out["scoring_roofline"] = roofline_entry(
        " + ".join(("k_mx_coefs + k_mx_reject + k_mx_exact<%d> (%s scoring, FP16 hi/lo rejection on the matrix "
                    "cores, FP64 exact pass)" if mx_on else "k_score<%d> (%s scoring)")
                   % (1 if k else 0, "Sampson" if k else "transfer") for k in kinds_run),
        "fp64-valu", score_ach, sc_launch, sc_ms / max(1, sc_launch),
        note="algorithmic %s flop per hypothesis-point (FP64 equivalent), priced against the FP64 peak"
             % "/".join("%d" % FLOP_PER_HYP_PT[k] for k in kinds_run))
    if dominant == "fit" and fit_ms > 0:
        fit_ach = fit_flops / (fit_ms * 1e-3) / 1e12
        tr = TRAFFIC_KIB.get((name, pairs, groups))
        fit_points = sum(v["points_per_launch"] * v["launches_per_step"] for v in fit_parts.values()) * args.steps
        out["roofline"] = roofline_entry(
            "%s (non-minimal fits: LO, IRLS%s)" % (" + ".join(FIT_KERNEL[k] for k in kinds_run),
                                                   ", post-hoc BA" if 1 in kinds_run or 2 in kinds_run else ""),
            "fp64-valu", fit_ach, fit_launch, fit_ms / max(1, fit_launch),
            traffic=None if tr is None else 1024.0 * (tr[0] + tr[1]),
            algorithmic_bytes_per_launch=FIT_BYTES_PER_POINT * fit_points / max(1, fit_launch),
            flop_per_launch=fit_flops / max(1, fit_launch))
    else:
        out["roofline"] = dict(out["scoring_roofline"], traffic=None)
        if dominant in SOLVER_FAMILY:
            # a minimal solver (FP64 VALU, data-dependent branches): reported
            # as solves/s beside the scoring roofline, no roofline claim
            r = fams[dominant]
            out["roofline"]["note"] = ("the dominant family is the minimal solver %s, reported as solves/s in "
                                       "dominant_solver; the roofline above is the scoring kernels'"
                                       % SOLVER_FAMILY[dominant])
            out["dominant_solver"] = {"kernel": SOLVER_FAMILY[dominant], "ms_per_step": r[0] / args.steps,
                                      "launches_per_step": r[1] / args.steps,
                                      "solves_per_s": r[2] / (r[0] * 1e-3) if r[0] > 0 else 0.0}
    out["roofline"]["dominant_family"] = dominant
    if not args.no_cpu:
        cpu_scenes = make_pairs(name, 256, first=pairs + 1000)
        out["cpu_baseline"] = cpu_baseline(name, cpu_scenes, max_iters)
    runner.free()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
