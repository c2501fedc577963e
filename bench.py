"""Benchmark: search QPS at 10-recall@10 >= 0.95 and index build seconds on
one MI355X, reduced-precision traversal against FP32.

Headline workload: 10,000,000 x 200 float32, maximum inner product, a
5,000-component Gaussian mixture (centres N(0,1), scale 0.3, seed 8) with
10,000 queries from shifted centres (seed 9); Vamana R=64, L=128,
alpha=1.0; k=10.  The build is timed once (device time, vectors resident).
A step is one batched search launch over the 10,000 device-resident
queries.  For each traversal mode -- FP32 (the reference's arithmetic), BF16
and per-dimension-scaled FP8-e4m3 copies with FP32 accumulation and an FP32
re-rank of the final 100 -- the smallest beam of the reference's sweep list
(check_nn_recall.h:173-175) whose 10-recall@10 against FP64 brute-force
ground truth (K = 100, ties at the 10th distance counted as
check_nn_recall.h:86-109 counts them) is >= 0.95 is found, then W warm-up and
K timed steps run.  `value` is the QPS of the fastest reduced-precision mode;
every mode's QPS, beam, recall, kernel time and distance error against FP64
are in `modes`.

`configs` times the other workloads the same way (build seconds, QPS at
recall 0.95): 100,000 x 20 float32 L2 (with its per-query-beam launch),
1,000,000 x 100 int8 and 10,000,000 x 128 uint8.

`cpu_baseline` is the oracle (a CPU restatement of the reference's build and
search, g++ -O2 like its Makefile) on every core this process may use: the
FP32 search of all queries at the FP32 beam over the GPU-built graph (the
same graph gives the same visits), plus a build of the first --cpu-build-n
points as a sample of the build rate.  A full 10M-point CPU build does not
fit this run and is not extrapolated.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-configs]

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "search QPS at 10-recall@10=0.95 + index build sec, 1 MI355X, low-prec vs FP32"
HBM_PEAK = 8.0e12
BEAMS = [10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40, 45, 50, 55, 60, 65,
         70, 80, 90, 100, 120, 140, 160, 180, 200, 225, 250, 275, 300, 375, 500, 750, 1000]  # check_nn_recall.h:173-175
S_TRAV = {"bf16": 2, "fp8": 1}
GT_K = 100
RERANK = 100


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
        """(wall ms per step, kernel ms per step) over `steps` launches; the
        kernel time is the search kernel's own, from HIP events recorded on
        the stream it runs on."""
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
    (1 + degree words), every query its row, and a re-rank reads rerank_rows
    full FP32 rows per query."""
    return cmps * d * s_trav + vis * 4.0 * (1.0 + avg_deg) + nq * d * s_trav + nq * rerank_rows * d * 4.0


def sweep(ds, gt_ids, gt_d, trav="native", rerank=0, target=0.95):
    """Smallest beam of the reference's list reaching the target recall."""
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


def profiled_traffic(key):
    """HBM bytes per search launch measured by rocprofv3 FETCH_SIZE /
    WRITE_SIZE (copy-calibrated) for this exact workload, mode and beam, from
    a committed profile summary naming that key; (None, None) otherwise."""
    pdir = os.path.join(HERE, "profiles")
    if not os.path.isdir(pdir):
        return None, None
    for name in sorted(os.listdir(pdir)):
        path = os.path.join(pdir, name)
        key_ok, val = False, None
        for line in open(path):
            if line.startswith("bench_key:") and line.split(":", 1)[1].strip() == key:
                key_ok = True
            if line.startswith("traffic_bytes_per_launch:"):
                val = float(line.split(":")[1])
        if key_ok and val is not None:
            return val, "profiles/" + name
    return None, None


def cpu_info():
    model = platform.processor() or ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    quota = None  # a cgroup CPU quota caps the cores a process can use below its affinity mask
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = float(q) / float(period)
    except (OSError, ValueError):
        pass
    usable = affinity if quota is None else max(1, min(affinity, int(quota)))
    return {"model": model, "nproc": os.cpu_count(), "affinity_cores": affinity, "cgroup_cpu_quota": quota,
            "usable_cores": usable}


def measure_mode(ds, idx, X, Q, w, gi, gd, mode, steps, warmup, avg_deg, key):
    """Sweep, autotune and time one traversal mode of an index."""
    from distributed_diskann_amd.index import recall_with_ties
    rerank = 0 if mode == "native" else RERANK
    prep = idx.prepare_traversal(mode)
    Ls, r_sweep = sweep(ds, gi, gd, mode, rerank)
    tune = idx.autotune_search(Q, 10, Ls, trav=mode, rerank=rerank)
    wall, kern = ds.timed(Ls, steps, warmup, mode, rerank)
    rec = recall_with_ties(ds.result_ids(), gi, gd, 10)
    cm, vs = ds.counters()
    s_trav = (4 if w.dtype == "float32" else 1) if mode == "native" else S_TRAV[mode]
    nr = 0 if mode == "native" else min(RERANK, Ls)
    B = alg_bytes(cm, vs, ds.nq, w.d, s_trav, avg_deg, nr)
    out = {"Ls": Ls, "recall": rec, "qps": ds.nq / (wall / 1000.0), "ms_per_step": wall, "kernel_ms": kern,
           "alg_bytes_per_launch": B, "alg_GBps": B / (kern / 1000.0) / 1e9,
           "hbm_frac": B / (kern / 1000.0) / HBM_PEAK, "dist_cmps_per_query": cm / ds.nq,
           "visited_per_query": vs / ds.nq, "prepare_s": prep, "autotune": tune}
    if w.dtype == "float32":
        out["dist_rel_err_vs_fp64"] = distance_error(idx, X, Q, w.d, w.metric, mode)
    log(f"[bench] {key} {mode}: Ls={Ls} recall {rec:.4f} QPS {out['qps']:.0f} kernel {kern:.3f} ms "
        f"({out['alg_GBps']:.0f} GB/s algorithmic, {out['hbm_frac']:.3f} of HBM)")
    return out


def run_workload(torch, name, modes, steps, warmup, extra=None):
    """Generates, builds and times one workload; returns (result, X, Q, idx, graph)."""
    from distributed_diskann_amd import Index, datasets
    t0 = time.perf_counter()
    X, Q, w = datasets.generate(name)
    gen_s = time.perf_counter() - t0
    log(f"[bench] {name}: {X.shape[0]}x{w.d} {w.dtype} {w.metric} R={w.R} L={w.L} alpha={w.alpha} "
        f"nq={Q.shape[0]} (generated in {gen_s:.1f} s)")
    idx = Index(X, w.d, w.metric, w.R)
    build_s = idx.build(R=w.R, L=w.L, alpha=w.alpha)
    bprof = {k: v for k, v in idx.profile().items() if k.endswith("_ms")}
    G = idx.get_graph()
    avg_deg = float(G[:, 0].mean())
    log(f"[bench] {name}: build {build_s:.3f} s (device), average degree {avg_deg:.2f}")
    gi, gd = idx.ground_truth(Q, GT_K)
    gt_ms = idx.profile()["dense_ms"]
    log(f"[bench] {name}: FP64 ground truth top-{GT_K} {gt_ms:.1f} ms")
    ds = DeviceSearch(torch, idx, Q)
    res = {"workload": f"{X.shape[0]}x{w.d} {w.dtype} {w.metric}, R={w.R} L={w.L} alpha={w.alpha}, "
                       f"k=10, {Q.shape[0]} queries", "build_seconds": build_s, "build_profile_ms": bprof,
           "avg_degree": avg_deg, "ground_truth_ms": gt_ms, "ground_truth_k": GT_K, "modes": {}}
    for mode in modes:
        res["modes"][mode] = measure_mode(ds, idx, X, Q, w, gi, gd, mode, steps, warmup, avg_deg, name)
    if extra:
        extra(ds, idx, Q, gi, gd, res)
    del ds
    return res, X, Q, w, idx, G, gi, gd


def per_query_beams(torch, steps):
    def extra(ds, idx, Q, gi, gd, res):
        from distributed_diskann_amd import datasets
        from distributed_diskann_amd.index import recall_with_ties
        beams = datasets.s_query_beams(Q.shape[0])
        pqb = torch.from_numpy(beams.astype(np.int32)).to(torch.device("cuda:0"))
        idx.autotune_search(Q, 10, int(beams.max()))
        wall, kern = ds.timed(int(beams.max()), steps, 1, pqb=pqb, bmax=int(beams.max()))
        res["per_query_beam"] = {"qps": Q.shape[0] / (wall / 1000.0), "kernel_ms": kern,
                                 "recall": recall_with_ties(ds.result_ids(), gi, gd, 10),
                                 "beams": [int(b) for b in np.unique(beams)]}
        log(f"[bench] per-query beams: {res['per_query_beam']}")
    return extra


def cpu_baseline(X, Q, w, G, gi, gd, Ls, build_n):
    """The oracle on every usable core: FP32 search of all queries at beam Ls
    over the GPU-built graph, and a build of the first build_n points."""
    from distributed_diskann_amd.index import recall_with_ties
    from oracle import oracle as O
    info = cpu_info()
    threads = info["usable_cores"]
    m = min(build_n, X.shape[0])
    Ps = O.Points(np.ascontiguousarray(X[:m]), w.d, w.metric)
    t0 = time.perf_counter()
    O.build(Ps, w.R, w.L, w.alpha, tie_mode=O.TIE_REFERENCE, nthreads=threads)
    build_s = time.perf_counter() - t0
    del Ps
    P = O.Points(X, w.d, w.metric)
    Gg = O.Graph(G)
    O.search_all(P, Gg, Q[:200], 10, Ls, cut=1.35, nthreads=threads)  # warm-up (check_nn_recall.h:62-66)
    t0 = time.perf_counter()
    ci, _, _, _ = O.search_all(P, Gg, Q, 10, Ls, cut=1.35, nthreads=threads)
    search_s = time.perf_counter() - t0
    rec = recall_with_ties(ci, gi, gd, 10)
    log(f"[bench] CPU oracle ({threads} threads, {info['model']}): build of {m} points {build_s:.2f} s, "
        f"search {Q.shape[0] / search_s:.0f} QPS at Ls={Ls} (recall {rec:.4f})")
    return {"value": Q.shape[0] / search_s, "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"search of all {Q.shape[0]} queries at Ls={Ls} over the GPU-built "
                      f"{X.shape[0]}-point graph (recall {rec:.4f}); " +
                      (f"build of all {m} points in {build_s:.2f} s" if m == X.shape[0] else
                       f"build of the first {m} points in {build_s:.2f} s ({m / build_s:.0f} points/s); the full "
                       f"{X.shape[0]}-point CPU build does not fit this run and is not extrapolated"),
            "search_seconds": search_s, "build_points_sample": m, "build_seconds_sample": build_s,
            "cpu": info}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="headline workload only")
    ap.add_argument("--modes", default="native,bf16,fp8")
    ap.add_argument("--workload", default="mips", choices=["mips", "u8", "i8", "s"],
                    help="headline workload (mips: the 10M x 200 float32 set; the integer and 100K sets search "
                         "in their native arithmetic only; used for per-workload profiles)")
    ap.add_argument("--cpu-build-n", type=int, default=200_000)
    ap.add_argument("--cpu-configs", action="store_true",
                    help="also time the CPU oracle on the smaller configurations (and run the 10K x 20 one): "
                         "their sub-blocks then carry a cpu_baseline")
    ap.add_argument("--only-gt", action="store_true",
                    help="only time the FP64 / int8 MFMA ground truth (top-100) of the 10M x 200 float32 and "
                         "10M x 128 uint8 workloads and exit (for matrix-core counter profiles)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (calibrates the FETCH_SIZE / "
                         "WRITE_SIZE counters against known bytes) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported: this build runs on one "
              "GPU, and its multi-GPU path is a design, not code", file=sys.stderr)
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

    if args.only_gt:
        from distributed_diskann_amd import Index, datasets
        res = {}
        for name in ("mips", "u8"):
            X, Q, w = datasets.generate(name)
            idx = Index(X, w.d, w.metric, 8)
            for k in (10, GT_K):
                idx.ground_truth(Q, k)
                ms = idx.profile()["dense_ms"]
                flops = 2.0 * X.shape[0] * Q.shape[0] * w.d
                res[f"{name}_k{k}"] = {"ms": ms, "T_ops_per_s": flops / (ms / 1000.0) / 1e12}
                log(f"[bench] ground truth {name} k={k}: {ms:.1f} ms, {flops / (ms / 1000.0) / 1e12:.1f} T ops/s")
            idx.close()
            del X, Q
        print(json.dumps({"ground_truth": res}), flush=True)
        return

    modes = args.modes.split(",") if args.workload in ("mips", "s") else ["native"]
    head, X, Q, w, idx, G, gi, gd = run_workload(torch, args.workload, modes, args.steps, args.warmup)
    low = {m: r for m, r in head["modes"].items() if m != "native" and r["recall"] >= 0.95}
    best = max(low, key=lambda m: low[m]["qps"]) if low else ("native" if "native" in head["modes"] else modes[0])
    top = head["modes"][best]
    fp32 = head["modes"].get("native")
    X_n = X.shape[0]
    # the PCIe-inclusive rate: the same search through ddg_search with the
    # queries and results in host memory (copied in and out around the
    # launch); reported next to the device-resident value, never as it
    rr = 0 if best == "native" else RERANK
    idx.search(Q, 10, top["Ls"], 1.35, trav=best, rerank=rr)
    t0 = time.perf_counter()
    for _ in range(3):
        idx.search(Q, 10, top["Ls"], 1.35, trav=best, rerank=rr)
    host_ms = (time.perf_counter() - t0) * 1000.0 / 3
    host = {"qps": Q.shape[0] / (host_ms / 1000.0), "ms_per_call": host_ms,
            "note": "ddg_search with host query and result buffers (PCIe copies included)"}
    log(f"[bench] {args.workload} {best}: host buffers {host['qps']:.0f} QPS ({host_ms:.3f} ms per call)")
    idx.close()
    del idx
    torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(X, Q, w, G, gi, gd, fp32["Ls"] if fp32 else top["Ls"], args.cpu_build_n)
    del X, G

    configs = {}
    if not args.no_configs and args.workload == "mips":
        sub = [("s", per_query_beams(torch, max(1, args.steps // 4))), ("i8", None), ("u8", None)]
        if args.cpu_configs:
            sub.insert(0, ("xs", None))
        for name, extra in sub:
            r, Xc, Qc, wc, idxc, Gc, gic, gdc = run_workload(torch, name, ["native"], max(1, args.steps // 2), 1, extra)
            idxc.close()
            del idxc
            torch.cuda.empty_cache()
            if args.cpu_configs:
                r["cpu_baseline"] = cpu_baseline(Xc, Qc, wc, Gc, gic, gdc, r["modes"]["native"]["Ls"], args.cpu_build_n)
            del Xc, Qc, Gc
            configs[{"xs": "10Kx20_f32", "s": "100Kx20_f32", "i8": "1Mx100_i8", "u8": "10Mx128_u8"}[name]] = r

    key = f"{args.workload} {best} Ls={top['Ls']}"
    traffic, traffic_src = profiled_traffic(key)
    kname = f"k_search<{w.dtype},{w.metric},{best}>"
    out = {
        "metric": METRIC, "value": top["qps"], "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": top["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": ({"float32": "f32", "uint8": "u8 (int32 accumulate)", "int8": "i8 (int32 accumulate)"}[w.dtype]
                  if best == "native" else
                  f"{'bf16' if best == 'bf16' else 'fp8e4m3'} rows, f32 accumulate, f32 re-rank"),
        "data": f"synthetic: {head['workload']}; base seed {w.base_seed}, query seed {w.query_seed}",
        "config": {"workload": f"{head['workload']}, {best} traversal" +
                               (f" + FP32 re-rank of {RERANK}" if best != "native" else "") + f", Ls={top['Ls']}",
                   "n": int(X_n), "d": w.d, "nq": int(Q.shape[0]), "R": w.R, "L": w.L, "alpha": w.alpha, "k": 10,
                   "mode": best, "Ls": top["Ls"]},
        "recall": top["recall"],
        "build_seconds": head["build_seconds"],
        "fp32": None if fp32 is None else {"qps": fp32["qps"], "Ls": fp32["Ls"], "recall": fp32["recall"]},
        "speedup_vs_fp32": None if fp32 is None else top["qps"] / fp32["qps"],
        "host_buffers": host,
        "timing": ("value and ms_per_step: whole search calls, wall clock, inputs and outputs in device memory; "
                   "from a beam of 140 on a call also runs a short pre-pass search (beam 10, 8 visits) and a "
                   "key sort that order its queries (about 0.35 ms). kernel_ms and roofline: the main search "
                   "launch alone, HIP events on its stream"),
        "headline": head,
        "configs": configs,
        "roofline": {"bound": "hbm", "achieved": top["alg_GBps"], "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": top["hbm_frac"], "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": kname, "kernel_avg_ms": top["kernel_ms"],
                     "alg_bytes_per_launch": top["alg_bytes_per_launch"],
                     "note": (f"{X_n} rows: the traversed copy and the graph are far above the 256 MB Infinity "
                              "Cache" if X_n >= 1_000_000 else "the working set is cache-resident")},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
