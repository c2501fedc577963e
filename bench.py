"""Benchmark: search QPS at 10-recall@10 >= 0.95 and index build seconds on
one MI355X.

Main line (`value`): the 100,000 x 20 float32 workload (seed 2 base, seed 3
queries; R=32, L=64, alpha=1.2; k=10).  A step is one batched search of the
10,000 device-resident queries at the smallest uniform beam width whose
10-recall@10 (FP64 brute-force ground truth, ties counted as in
check_nn_recall.h:86-109) is >= 0.95.  The build is timed once (device time,
vectors resident), and the per-query-beam launch of the workload (beams
uniform over {10, 20, 50, 100, 200}) is timed too.

`low_precision`: the maximum-inner-product mixture workload (5,000 centres
N(0,1), scale 0.3, seed 8; queries from shifted centres, seed 9; d=200,
R=64, L=128, alpha=1.0), at --lowp-n base points: QPS at recall 0.95 for
FP32, BF16 and FP8-e4m3 traversal (re-rank of the final 100 in FP32), with
each mode's distance error against FP64.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--lowp-n N]
    python bench.py --workload {i8,u8,mips} [--n N] [--nq NQ] [--lowp-n 0]

Other workloads run the same protocol (build, FP64 ground truth, Ls sweep,
timed device-resident search; float32 workloads also time BF16 and FP8
traversal with FP32 re-rank).  Above --cpu-full-max points the CPU baseline
is a bounded sample: the oracle builds the first --cpu-build-n points, and
searches --cpu-nq queries over the GPU-built graph at the chosen Ls.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "search QPS at 10-recall@10=0.95 + index build sec, 1 MI355X, low-prec vs FP32"
HBM_PEAK = 8.0e12
BEAMS = [10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40, 45, 50, 55, 60, 65,
         70, 80, 90, 100, 120, 140, 160, 180, 200, 225, 250, 275, 300, 375, 500, 750, 1000]  # check_nn_recall.h:173-175
ESIZE = {"native": None, "bf16": 2, "fp8": 1}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class DeviceSearch:
    """Device-resident queries and outputs; one call = one search launch."""

    def __init__(self, torch, idx, Q, k=10):
        dev = torch.device("cuda:0")
        self.torch, self.idx, self.k = torch, idx, k
        self.nq = Q.shape[0]
        self.Qd = torch.from_numpy(Q).to(dev)
        self.starts = torch.zeros(1, dtype=torch.int32, device=dev)
        self.ids = torch.empty((self.nq, k), dtype=torch.int32, device=dev)
        self.d = torch.empty((self.nq, k), dtype=torch.float32, device=dev)
        self.cm = torch.empty(self.nq, dtype=torch.int64, device=dev)
        self.vs = torch.empty(self.nq, dtype=torch.int32, device=dev)

    def run(self, Ls, trav="native", rerank=0, pqb=None, bmax=None):
        self.idx.search_device(self.Qd, self.k, Ls, self.ids, self.d, self.cm, self.vs, cut=1.35,
                               starts=self.starts, per_query_beam=pqb, beam_max=bmax, trav=trav, rerank=rerank)

    def timed(self, Ls, steps, warmup, trav="native", rerank=0, pqb=None, bmax=None):
        """(wall ms per step, kernel ms per step) over `steps` launches."""
        for _ in range(warmup):
            self.run(Ls, trav, rerank, pqb, bmax)
        self.torch.cuda.synchronize()
        kern = 0.0
        t0 = time.perf_counter()
        for _ in range(steps):
            before = self.idx.profile()["search_ms"]
            self.run(Ls, trav, rerank, pqb, bmax)
            kern += self.idx.profile()["search_ms"] - before
        self.torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1000.0 / steps, kern / steps

    def result_ids(self):
        return self.ids.cpu().numpy().view(np.uint32)

    def counters(self):
        return float(self.cm.sum().item()), float(self.vs.sum().item())


def alg_bytes(cmps, vis, nq, d, s_trav, avg_deg, rerank_rows=0):
    """Algorithmic bytes of one search launch: every distance reads one
    d-element row of the traversed copy, every expansion one adjacency row
    (1 + degree words), every query its FP32/native row, and a re-rank reads
    rerank_rows full FP32 rows per query."""
    return cmps * d * s_trav + vis * 4.0 * (1.0 + avg_deg) + nq * d * s_trav + nq * rerank_rows * d * 4.0


def sweep(ds, gt_ids, gt_d, trav="native", rerank=0, target=0.95):
    from distributed_diskann_amd.index import recall_with_ties
    last = None
    for Ls in BEAMS:
        ds.run(Ls, trav, rerank)
        r = recall_with_ties(ds.result_ids(), gt_ids, gt_d, 10)
        last = (Ls, r)
        if r >= target:
            return Ls, r
    return last


def distance_error(idx, X, Q, d, metric, mode, nq=100, m=1000, seed=17):
    """Relative error of `mode` traversal distances against FP64, normalised by
    |q| |x| (scale-free for both metrics): median, p99, max."""
    rng = np.random.default_rng(seed)
    qi = rng.integers(0, Q.shape[0], nq)
    ids = rng.integers(0, X.shape[0], (nq, m)).astype(np.uint32)
    got = idx.distances(Q[qi], ids, mode).astype(np.float64)
    err = np.empty((nq, m))
    for i in range(nq):
        q = Q[qi[i], :d].astype(np.float64)
        xs = X[ids[i], :d].astype(np.float64)
        ex = ((xs - q) ** 2).sum(1) if metric == "l2" else -(xs @ q)
        err[i] = np.abs(got[i] - ex) / (np.linalg.norm(q) * np.linalg.norm(xs, axis=1) + 1e-30)
    return {"median": float(np.median(err)), "p99": float(np.quantile(err, 0.99)), "max": float(err.max())}


def profiled_traffic(workload, n, Ls):
    """HBM bytes per search launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE
    (copy-calibrated) for this exact workload and beam, from the committed
    profile summary; None when no profile covers it."""
    covered = {("s", 100_000, 80): "k_search_f32_100Kx20_Ls80.md",
               ("u8", 10_000_000, 140): "k_search_u8_10Mx128_Ls140.md"}
    name = covered.get((workload, n, Ls))
    if name is None or not os.path.exists(os.path.join(HERE, "profiles", name)):
        return None, None
    for line in open(os.path.join(HERE, "profiles", name)):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1]), "profiles/" + name
    return None, None


def low_precision(torch, args):
    from distributed_diskann_amd import Index, datasets
    X, Q, w = datasets.generate("mips", n=args.lowp_n, nq=args.lowp_nq)
    nq = Q.shape[0]
    log(f"[bench] low-precision workload: n={X.shape[0]} d={w.d} mips R={w.R} L={w.L} alpha={w.alpha} nq={nq}")
    idx = Index(X, w.d, "mips", w.R)
    build_s = idx.build(R=w.R, L=w.L, alpha=w.alpha)
    G = idx.get_graph()
    avg_deg = float(G[:, 0].mean())
    del G
    gi, gd = idx.ground_truth(Q, 10)
    gt_ms = idx.profile()["dense_ms"]
    log(f"[bench] lowp build {build_s:.2f} s, avg degree {avg_deg:.2f}, FP64 GT {gt_ms:.0f} ms")
    ds = DeviceSearch(torch, idx, Q)
    out = {"workload": f"{X.shape[0]}x{w.d} f32 MIPS mixture, R={w.R} L={w.L} alpha={w.alpha}, k=10, nq={nq}",
           "build_seconds": build_s, "gt_ms": gt_ms, "modes": {}}
    for mode in ("native", "bf16", "fp8"):
        rerank = 0 if mode == "native" else 100
        prep = idx.prepare_traversal(mode)
        Ls, rec = sweep(ds, gi, gd, mode, rerank)
        idx.autotune_search(Q, 10, Ls, trav=mode, rerank=rerank)
        wall, kern = ds.timed(Ls, max(3, args.steps // 4), 1, mode, rerank)
        cm, vs = ds.counters()
        s_trav = 4 if mode == "native" else ESIZE[mode]
        nr = 0 if mode == "native" else min(100, Ls)
        B = alg_bytes(cm, vs, nq, w.d, s_trav, avg_deg, nr)
        err = distance_error(idx, X, Q, w.d, "mips", mode) if mode != "native" else None
        out["modes"][mode] = {"Ls": Ls, "recall": rec, "qps": nq / (wall / 1000.0), "kernel_ms": kern,
                              "wall_ms": wall, "alg_GBps": B / (kern / 1000.0) / 1e9,
                              "hbm_frac": B / (kern / 1000.0) / HBM_PEAK, "prepare_s": prep,
                              "dist_rel_err_vs_fp64": err}
        log(f"[bench] lowp {mode}: Ls={Ls} recall {rec:.4f} QPS {nq / (wall / 1000.0):.0f} kernel {kern:.3f} ms "
            f"{B / (kern / 1000.0) / 1e9:.0f} GB/s err {err}")
    del ds, idx
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="s")
    ap.add_argument("--n", type=int, default=0, help="base points (0 = the workload's own size)")
    ap.add_argument("--nq", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-full-max", type=int, default=200_000)
    ap.add_argument("--cpu-build-n", type=int, default=100_000)
    ap.add_argument("--cpu-nq", type=int, default=1000)
    ap.add_argument("--lowp-n", type=int, default=1_000_000, help="base points of the low-precision workload "
                                                                  "(0 skips it)")
    ap.add_argument("--lowp-nq", type=int, default=10_000)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (calibrates the FETCH_SIZE / "
                         "WRITE_SIZE counters against known bytes) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the multi-GPU path is "
              "design-only)", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_probe:
        nbytes = args.copy_probe << 20
        src = torch.ones(nbytes, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": nbytes, "launches": 5}), flush=True)
        return

    from distributed_diskann_amd import Index, datasets
    from distributed_diskann_amd.index import recall_with_ties

    t0 = time.perf_counter()
    X, Q, w = datasets.generate(args.workload, n=args.n or None, nq=args.nq or None)
    n, nq = X.shape[0], Q.shape[0]
    log(f"[bench] workload {w.name}: n={n} d={w.d} {w.dtype} {w.metric} R={w.R} L={w.L} alpha={w.alpha} nq={nq} "
        f"(generated in {time.perf_counter() - t0:.1f} s)")
    idx = Index(X, w.d, w.metric, w.R)
    t0 = time.perf_counter()
    build_s = idx.build(R=w.R, L=w.L, alpha=w.alpha)
    build_wall = time.perf_counter() - t0
    bprof = idx.profile()
    G = idx.get_graph()
    avg_deg = float(G[:, 0].mean())
    log(f"[bench] build {build_s:.3f} s device ({build_wall:.3f} s wall); avg degree {avg_deg:.2f}")

    gt_ids, gt_d = idx.ground_truth(Q, 10)
    gt_ms = idx.profile()["dense_ms"]
    log(f"[bench] FP64 ground truth {gt_ms:.1f} ms")

    ds = DeviceSearch(torch, idx, Q)
    chosen, rec0 = sweep(ds, gt_ids, gt_d)
    log(f"[bench] sweep -> Ls={chosen} (recall {rec0:.4f})")
    tune = idx.autotune_search(Q, 10, chosen)
    log(f"[bench] autotune at Ls={chosen}: {tune}")
    ms_per_step, kern_avg_ms = ds.timed(chosen, args.steps, args.warmup)
    qps = nq / (ms_per_step / 1000.0)
    rec = recall_with_ties(ds.result_ids(), gt_ids, gt_d, 10)
    tot_cmps, tot_vis = ds.counters()
    s = 4 if w.dtype == "float32" else 1
    achieved = alg_bytes(tot_cmps, tot_vis, nq, w.d, s, avg_deg) / (kern_avg_ms / 1000.0)
    log(f"[bench] Ls={chosen}: {ms_per_step:.3f} ms/step wall, kernel {kern_avg_ms:.3f} ms, QPS {qps:.0f}, "
        f"recall {rec:.4f}, {achieved / 1e9:.1f} GB/s algorithmic")

    # the per-query-beam launch of the workload
    beams = datasets.s_query_beams(nq)
    pqb = torch.from_numpy(beams.astype(np.int32)).to(torch.device("cuda:0"))
    idx.autotune_search(Q, 10, int(beams.max()))  # the launch shape of the largest beam class
    mixed_ms, mixed_kern = ds.timed(chosen, max(1, args.steps // 4), 1, pqb=pqb, bmax=int(beams.max()))
    mixed_rec = recall_with_ties(ds.result_ids(), gt_ids, gt_d, 10)
    log(f"[bench] per-query beams: {mixed_ms:.3f} ms/step, recall {mixed_rec:.4f}")

    # reduced-precision traversal of the same index (float32 workloads)
    modes = None
    if w.dtype == "float32":
        modes = {}
        for mode in ("bf16", "fp8"):
            prep = idx.prepare_traversal(mode)
            Ls_m, r_m = sweep(ds, gt_ids, gt_d, mode, 100)
            idx.autotune_search(Q, 10, Ls_m, trav=mode, rerank=100)
            wall_m, kern_m = ds.timed(Ls_m, max(3, args.steps // 4), 1, mode, 100)
            cm_m, vs_m = ds.counters()
            B_m = alg_bytes(cm_m, vs_m, nq, w.d, ESIZE[mode], avg_deg, min(100, Ls_m))
            modes[mode] = {"Ls": Ls_m, "recall": r_m, "qps": nq / (wall_m / 1000.0), "kernel_ms": kern_m,
                           "alg_GBps": B_m / (kern_m / 1000.0) / 1e9, "prepare_s": prep}
            log(f"[bench] {mode}: Ls={Ls_m} recall {r_m:.4f} QPS {nq / (wall_m / 1000.0):.0f} kernel {kern_m:.3f} ms")
    del ds

    lowp = low_precision(torch, args) if (args.lowp_n > 0 and args.workload == "s") else None

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        if n <= args.cpu_full_max:
            P = O.Points(X, w.d, w.metric)
            t0 = time.perf_counter()
            Gc = O.build(P, w.R, w.L, w.alpha, tie_mode=O.TIE_REFERENCE, nthreads=threads)
            cpu_build = time.perf_counter() - t0
            O.search_all(P, Gc, Q, 10, chosen, cut=1.35, nthreads=threads)  # warm-up (check_nn_recall.h:62-66)
            t0 = time.perf_counter()
            ci, _, _, _ = O.search_all(P, Gc, Q, 10, chosen, cut=1.35, nthreads=threads)
            cpu_search = time.perf_counter() - t0
            cpu_rec = recall_with_ties(ci, gt_ids, gt_d, 10)
            cpu = {"value": nq / cpu_search, "unit": "queries/s", "cores": threads, "kind": "port",
                   "sample": f"full workload: build of all {n} points and {nq} queries at Ls={chosen} "
                             f"(CPU recall {cpu_rec:.4f})",
                   "build_seconds": cpu_build}
            log(f"[bench] CPU ({threads} threads): build {cpu_build:.2f} s, QPS {nq / cpu_search:.0f}")
        else:
            # bounded sample: a build of the first cpu_build_n points, and cpu_nq
            # queries searched over the GPU-built graph (the same graph, so the
            # same visits) at the chosen Ls
            m = min(args.cpu_build_n, n)
            Ps = O.Points(np.ascontiguousarray(X[:m]), w.d, w.metric)
            t0 = time.perf_counter()
            O.build(Ps, w.R, w.L, w.alpha, tie_mode=O.TIE_REFERENCE, nthreads=threads)
            cpu_build_m = time.perf_counter() - t0
            del Ps
            P = O.Points(X, w.d, w.metric)
            Gg = O.Graph(G)
            qs = np.ascontiguousarray(Q[:args.cpu_nq])
            O.search_all(P, Gg, qs[:100], 10, chosen, cut=1.35, nthreads=threads)  # warm-up
            t0 = time.perf_counter()
            O.search_all(P, Gg, qs, 10, chosen, cut=1.35, nthreads=threads)
            cpu_search = time.perf_counter() - t0
            cpu = {"value": qs.shape[0] / cpu_search, "unit": "queries/s", "cores": threads, "kind": "port",
                   "sample": f"{qs.shape[0]} of {nq} queries at Ls={chosen} over the GPU-built {n}-point graph; "
                             f"build of the first {m} points only",
                   "build_seconds_sample": cpu_build_m, "build_points_per_s_sample": m / cpu_build_m}
            log(f"[bench] CPU ({threads} threads): build of {m} points {cpu_build_m:.2f} s "
                f"({m / cpu_build_m:.0f} points/s), QPS {qs.shape[0] / cpu_search:.0f}")
            del P, Gg

    traffic, traffic_src = profiled_traffic(args.workload, n, chosen)
    out = {
        "metric": METRIC, "value": qps, "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": {"float32": "f32", "uint8": "u8", "int8": "i8"}[w.dtype],
        "data": f"synthetic {w.name}: {n}x{w.d} {w.dtype} seed {w.base_seed}, {nq} queries seed {w.query_seed}",
        "config": {"workload": f"{n}x{w.d} {w.dtype} {w.metric}, R={w.R} L={w.L} alpha={w.alpha}, k=10, Ls={chosen}",
                   "n": n, "d": w.d, "nq": nq, "R": w.R, "L": w.L, "alpha": w.alpha, "k": 10, "Ls": chosen},
        "recall": rec,
        "autotune": tune,
        "build_seconds": build_s,
        "build_profile_ms": {k: v for k, v in bprof.items() if k.endswith("_ms")},
        "per_query_beam": {"qps": nq / (mixed_ms / 1000.0), "kernel_ms": mixed_kern, "recall": mixed_rec,
                           "beams": [10, 20, 50, 100, 200]},
        "reduced_traversal": modes,
        "low_precision": lowp,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": f"k_search<{w.dtype},{w.metric},native>", "kernel_avg_ms": kern_avg_ms,
                     "note": (f"{X.nbytes / 1e6:.1f} MB of vectors" +
                              (": the working set is cache-resident" if X.nbytes < 256e6 else ""))},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
