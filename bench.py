#!/usr/bin/env python3
"""Benchmark: image pairs/s of Graph-Cut RANSAC on one MI355X.

Workload (default): fundamental-matrix GC-RANSAC on synthetic two-view scenes
-- f=800 pinhole cameras, 640x480, 5,000 correspondences, 30% inliers with
0.5 px noise, 1 px Sampson threshold, uniform sampler, grid neighbourhood
(8 divisions), spatial coherence weight 0.975, confidence 0.99, max_iters
10^6 so every pair runs to the 0.99-confidence stopping rule.  A step is one
engine run over a batch of `--pairs` independent pairs (default 1,024, the
batch size of the mixed workload) whose correspondences are resident in HBM
before the timed region.

Reported next to the throughput:
  * roofline of the dominant kernel family by HIP-event time on the engine
    stream over the timed steps.  For both workloads that is the
    non-minimal fit kernel (k_fit_f / k_fit_h): achieved = the algorithmic
    FP64 flops of the fits (the engine's count: per-point costs of the
    normalisation, the DLT / eight-point systems, the QR and every LM pass
    the device reports) over the kernel time, against the FP64 peak;
  * the scoring kernel's roofline (42 / 18 flop per hypothesis-point);
  * hypothesis-points/s of the scoring kernel;
  * the plain path (one pair per engine run) pairs/s;
  * cpu_baseline: the C++ oracle restatement of the reference on one host
    core over a bounded sample of the same pairs.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--pairs B]
                       [--workload fundamental|homography|guided|mixed] [--no-cpu]

--workload guided: 1,024 held-out fundamental pairs of 4,000 correspondences at
5 % inliers (descriptor-ratio scores as the synthetic generator defines them),
solved by the learned-guidance path with the committed network
(pygcransac/data/guided_mlp.npz, written by pygcransac.train_guided); the
plain path beside it runs the uniform sampler one pair per call.

--workload mixed: half homography (3 px) and half
fundamental (1 px), correspondence counts log-uniform in [200, 8000] and
inlier ratios uniform in [0.1, 0.7], max_iters 10^4 (the API default).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "graph-cut-ransac_amd"))

import numpy as np  # noqa: E402

from pygcransac import _native, hipmem, synthetic  # noqa: E402

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64

FIT_KERNEL = {0: "k_fit_h", 1: "k_fit_f"}

# HBM traffic per launch of the fit kernel (KiB read, KiB written), from
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of this script's workloads
# (one batched step of the default 1,024 pairs), scaled by the counters'
# reading for a 256 MiB device copy (FETCH_SIZE x2.0, WRITE_SIZE x1.0); the
# summary with the commands is profiles/pmc_traffic_fit_kernels.md.
TRAFFIC_KIB = {
    "fundamental": (8572.6, 149.6),  # k_fit_f<64> and k_fit_f<256>, averaged over their launches
    "homography": (8342.1, 171.2),  # k_fit_h<64> and k_fit_h<256>
}
TRAFFIC_PAIRS = 1024  # the batch size those counters were read at

WORKLOADS = {
    # name: (kind, generator, n, inlier_ratio, noise, threshold, image (w, h), config id, flop per hyp-pt)
    "fundamental": (1, synthetic.fundamental_scene, 5000, 0.3, 0.5, 1.0, (640, 480), 1, 42.0),
    "homography": (0, synthetic.homography_scene, 2000, 0.5, 0.5, 3.0, (640, 480), 0, 18.0),
    # held-out pairs of the learned-guidance evaluation: 4,000 correspondences at 5 % inliers
    "guided": (1, synthetic.fundamental_scene, 4000, 0.05, 0.5, 1.0, (640, 480), 4, 42.0),
    # mixed batch: half homography pairs (threshold 3 px), half fundamental
    # pairs (1 px), correspondence counts log-uniform in [200, 8000], inlier
    # ratios uniform in [0.1, 0.7]; kind, n and ratio are drawn per pair
    "mixed": (-1, None, 0, 0.0, 0.5, None, (640, 480), 3, None),
}

# max_iters per workload: the fundamental and homography pairs run to the
# 0.99-confidence rule; at 5 % inliers the rule asks for ~6e9 seven-point
# samples, so the guided workload runs at the API default cap of 10^4, and so
# does the mixed batch (its fundamental pairs at 10 % inliers would ask for
# ~5e7); --max-iters overrides
MAX_ITERS = {"fundamental": 1000000, "homography": 1000000, "guided": 10000, "mixed": 10000}
THRESHOLD = {0: 3.0, 1: 1.0}
FLOP_PER_HYP_PT = {0: 18.0, 1: 42.0}


def make_pairs(name, count, first=0):
    kind, gen, n, eps, noise, thr, (w, h), cid, _ = WORKLOADS[name]
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


def params_for(name, seed, kind, thr, sampler=0):
    p = _native.GcrParams()
    _native.lib().gcr_default_params(ctypes.byref(p), kind)
    p.threshold = thr
    p.confidence = 0.99
    p.spatial_coherence_weight = 0.975
    p.max_iters = MAX_ITERS[name]
    p.min_iters = 50
    p.sampler = sampler
    p.neighborhood = 0
    p.neighborhood_size = 8.0
    p.lo_number = 50
    p.seed = seed
    return p


class Batch:
    def __init__(self, name, scenes):
        kind, _, _, _, _, _, (w, h), _, _ = WORKLOADS[name]
        self.n_pairs = len(scenes)
        corr = np.ascontiguousarray(np.concatenate([s["corr"] for s in scenes]))
        self.offsets = np.zeros(self.n_pairs + 1, np.int64)
        self.offsets[1:] = np.cumsum([len(s["corr"]) for s in scenes])
        self.kinds = np.array([s["kind"] for s in scenes], np.int32)
        self.sizes = np.tile(np.array([h, w, h, w], np.int32), self.n_pairs)
        self.params = (_native.GcrParams * self.n_pairs)()
        for i in range(self.n_pairs):
            self.params[i] = params_for(name, i, scenes[i]["kind"], scenes[i]["thr"])
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
            ctypes.c_void_p(self.dev.ptr.value), 1, self.sizes[4 * lo:4 * hi].ctypes.data_as(vp), None,
            ctypes.cast(ctypes.byref(self.params, lo * ctypes.sizeof(_native.GcrParams)), vp),
            self.models[9 * lo:].ctypes.data_as(vp),
            self.masks[int(self.offsets[lo]):].ctypes.data_as(vp), self.ninl[lo:].ctypes.data_as(vp),
            ctypes.cast(ctypes.byref(self.stats, lo * ctypes.sizeof(_native.GcrStats)), vp))
        _native.raise_for(rc)


class GuidedBatch:
    """The learned-guidance path over the same pairs: features, network,
    per-pair sort and PROSAC GC-RANSAC in one gcr_find_batch_guided call,
    with correspondences and descriptor-ratio scores resident in HBM."""

    def __init__(self, name, scenes, net):
        self.plain = Batch(name, scenes)
        for i in range(self.plain.n_pairs):
            self.plain.params[i] = params_for(name, i, scenes[i]["kind"], scenes[i]["thr"], sampler=1)
        ratio = np.ascontiguousarray(np.concatenate([synthetic.descriptor_ratio(s["seed"], s["gt"])
                                                     for s in scenes]))
        self.dev_ratio = hipmem.DeviceArray(ratio)
        self.net = np.ascontiguousarray(net, np.float32)

    def run(self):
        b = self.plain
        vp = ctypes.c_void_p
        rc = _native.lib().gcr_find_batch_guided(
            b.n_pairs, b.kinds.ctypes.data_as(vp), b.offsets.ctypes.data_as(vp), ctypes.c_void_p(b.dev.ptr.value),
            1, b.sizes.ctypes.data_as(vp), None, ctypes.c_void_p(self.dev_ratio.ptr.value),
            self.net.ctypes.data_as(vp), ctypes.cast(b.params, vp), b.models.ctypes.data_as(vp),
            b.masks.ctypes.data_as(vp), b.ninl.ctypes.data_as(vp), ctypes.cast(b.stats, vp), None)
        _native.raise_for(rc)

    def free(self):
        self.dev_ratio.free()


def recall(batch, scenes):
    """Per pair: fraction of the generator's true inliers in the returned mask."""
    out = []
    for i, s in enumerate(scenes):
        m = batch.masks[int(batch.offsets[i]):int(batch.offsets[i + 1])].astype(bool)
        out.append((m & s["gt"]).sum() / max(1, s["gt"].sum()))
    return np.array(out)


def cpu_baseline(name, scenes, budget_s):
    """The C++ oracle restatement on one host thread, cycling over `scenes`
    (each pass with fresh RANSAC seeds) until `budget_s` of oracle time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py  # the C++ oracle restatement (test infrastructure)

    _, _, _, _, _, _, (w, h), _, _ = WORKLOADS[name]
    spent = 0.0
    done = 0
    while spent < budget_s or done < 2:
        sc = scenes[done % len(scenes)]
        p = oracle_py.params(sc["thr"], 0.99, 0.975, MAX_ITERS[name], 50, 0, 50, 8.0, done)
        t0 = time.perf_counter()
        if sc["kind"] == 1:
            oracle_py.find_fundamental(sc["corr"], h, w, h, w, p)
        else:
            oracle_py.find_homography(sc["corr"], h, w, h, w, p)
        spent += time.perf_counter() - t0
        done += 1
    return {"value": done / spent, "unit": "pairs/s", "cores": 1, "kind": "port",
            "sample": "%d runs over %d %s pairs of the benchmark workload (seed = run index), one host thread, "
                      "%.1f s" % (done, min(done, len(scenes)), name, spent)}


def copy_calibration(mib):
    """One device-to-device copy of `mib` MiB (a plain copy kernel of known
    size), used to scale HBM byte counters read by a profiler."""
    nbytes = mib << 20
    src = hipmem.DeviceBuffer(nbytes)
    dst = hipmem.DeviceBuffer(nbytes)
    hipmem.copy_d2d(dst, src, nbytes)
    hipmem.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes}))
    src.free()
    dst.free()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=1024, help="pairs per step")
    ap.add_argument("--workload", default="fundamental", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--max-iters", type=int, default=None, help="override the workload's max_iters")
    ap.add_argument("--net", default=None, help="network parameter file of the guided workload "
                    "(default: the committed pygcransac/data/guided_mlp.npz)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--plain-pairs", type=int, default=4)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run one device copy of this many MiB (profiler counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box in this setup has one MI355X; "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration_mib > 0:
        copy_calibration(args.copy_calibration_mib)
        return
    L = _native.lib()
    if L.gcr_device_count() <= 0:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    name = args.workload
    kind, _, n, eps, noise, thr, (w, h), cid, flop = WORKLOADS[name]
    if args.max_iters:
        MAX_ITERS[name] = args.max_iters
    scenes = make_pairs(name, args.pairs)
    if name == "mixed":
        n = sum(len(s["corr"]) for s in scenes)
    batch = Batch(name, scenes)
    runner = batch
    if name == "guided":
        from pygcransac import guided
        runner = GuidedBatch(name, scenes, guided.load(args.net) if args.net else guided.load())
    print("# workload %s: %d pairs, %d correspondences %s per step" % (name, args.pairs, n,
                                                                      "in all" if name == "mixed" else "per pair"),
          file=sys.stderr, flush=True)
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
    host = {}
    for ph in ("copies", "chunk", "label", "fit", "masklist", "resume"):
        r = _native.profile_read("host_" + ph)
        if r:
            host[ph] = round(r[0] / args.steps, 3)
    kern = {}
    fams = {}
    for fam_name in ("solve_h4", "solve_f7", "solve_e5", "score_h", "score_f", "score_e", "energy", "mincut", "fit"):
        r = _native.profile_read(fam_name)
        if r:
            kern[fam_name] = round(r[0] / args.steps, 3)
            fams[fam_name] = r
    fit_parts = {}
    for sub in ("fit_lo", "fit_irls", "fit_post"):
        r = _native.profile_read(sub)
        if r:
            fit_parts[sub] = {"ms_per_step": round(r[0] / args.steps, 3), "launches_per_step": r[1] / args.steps,
                              "avg_launch_ms": r[0] / max(1, r[1]), "points_per_launch": r[2] / max(1, r[1])}
    kinds_run = sorted(set(int(s["kind"]) for s in scenes))
    sc_parts = [(k, _native.profile_read("score_f" if k == 1 else "score_h")) for k in kinds_run]
    sc_parts = [(k, r) for k, r in sc_parts if r]
    fam = "score_f" if kind == 1 else "score_h"
    sc_ms = sum(r[0] for _, r in sc_parts)
    sc_launch = sum(r[1] for _, r in sc_parts)
    sc_hp = sum(r[2] for _, r in sc_parts)
    sc_flop = sum(FLOP_PER_HYP_PT[k] * r[2] for k, r in sc_parts)
    fit_ms, fit_launch, fit_flops = _native.profile_read("fit") or (0.0, 0, 0.0)
    dominant = max(fams, key=lambda k: fams[k][0]) if fams else fam
    value = args.pairs * args.steps / dt
    res = runner.plain if runner is not batch else batch
    inl = np.array([res.stats[i].num_inliers for i in range(res.n_pairs)])
    iters = np.array([res.stats[i].iteration_number for i in range(res.n_pairs)])
    run_recall = recall(res, scenes)
    achieved = sc_flop / (sc_ms * 1e-3) / 1e12 if sc_ms > 0 else 0.0
    # plain path: one pair per engine run, same pairs
    plain = None
    if args.plain_pairs > 0:
        tp = time.perf_counter()
        for i in range(args.plain_pairs):
            batch.run(i, i + 1)
        hipmem.synchronize()
        plain = args.plain_pairs / (time.perf_counter() - tp)
        plain_recall = recall(batch, scenes)[:args.plain_pairs]
        plain_inl = np.array([batch.stats[i].num_inliers for i in range(args.plain_pairs)])
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
        "config": {
            "workload": ("mixed GC-RANSAC batch, %d pairs per step (half homography at 3 px, half fundamental at "
                         "1 px), %d correspondences in all (per pair log-uniform in [200, 8000]), inlier ratios "
                         "uniform in [0.1, 0.7], noise 0.5 px, grid 8 divisions, lambda 0.975, conf 0.99, "
                         "max_iters %d, uniform sampler" % (args.pairs, n, MAX_ITERS[name])) if name == "mixed" else
                        "%s GC-RANSAC, %d pairs per step, N=%d, inlier ratio %.2f, noise %.1f px, threshold %.1f px, "
                        "grid 8 divisions, lambda 0.975, conf 0.99, max_iters %d, %s" %
                        ("fundamental" if name == "guided" else name, args.pairs, n, eps, noise, thr, MAX_ITERS[name],
                         "PROSAC over the learned network's order (features, network and sort inside the timed "
                         "region)" if name == "guided" else "uniform sampler"),
            "pairs_per_step": args.pairs,
            "correspondences": n,
        },
        "plain_gpu_pairs_per_s": plain,
        "plain_gpu_path": "one pair per engine run, uniform sampler, same pairs and max_iters",
        "mean_true_inlier_recall": float(run_recall.mean()),
        "plain_gpu_mean_true_inlier_recall": float(plain_recall.mean()) if plain else None,
        "plain_gpu_mean_inliers": float(plain_inl.mean()) if plain else None,
        "mean_inliers_same_pairs_as_plain": float(inl[:args.plain_pairs].mean()) if plain else None,
        "host_ms_per_step": host,
        "kernel_ms_per_step": kern,
        "fit_kernel_breakdown": fit_parts,
        "scoring_hyp_pts_per_s": sc_hp / (sc_ms * 1e-3) if sc_ms > 0 else 0.0,
        "mean_iterations": float(iters.mean()),
        "mean_inliers": float(inl.mean()),
        "scoring_roofline": {
            "kernel": " + ".join("k_score<%d> (%s scoring)" % (k, "Sampson" if k == 1 else "transfer")
                                 for k in kinds_run),
            "bound": "mfma",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "avg_launch_ms": sc_ms / max(1, sc_launch),
            "launches": sc_launch,
        },
    }
    if dominant == "fit" and fit_ms > 0:
        fit_ach = fit_flops / (fit_ms * 1e-3) / 1e12
        tr = TRAFFIC_KIB.get(name)
        out["roofline"] = {
            "kernel": "%s (non-minimal fits: LO, IRLS%s)" % (" + ".join(FIT_KERNEL[k] for k in kinds_run),
                                                              ", post-hoc BA" if 1 in kinds_run else ""),
            "bound": "mfma",
            "achieved": fit_ach,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": fit_ach / FP64_PEAK_TFLOPS,
            "traffic": None if tr is None or args.pairs != TRAFFIC_PAIRS else 1024.0 * (tr[0] + tr[1]),
            "avg_launch_ms": fit_ms / max(1, fit_launch),
            "launches": fit_launch,
            "flop_per_launch": fit_flops / max(1, fit_launch),
        }
    else:
        out["roofline"] = dict(out["scoring_roofline"], traffic=None)
    out["roofline"]["dominant_family"] = dominant
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(name, make_pairs(name, {1: 16, 0: 64}.get(kind, 32), first=args.pairs),
                                           args.cpu_seconds)
    batch.dev.free()
    if runner is not batch:
        runner.free()
        runner.plain.dev.free()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
