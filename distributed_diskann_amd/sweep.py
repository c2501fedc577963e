"""The reference's recall / QPS reporting protocol over the GPU search.

search_and_parse (check_nn_recall.h:159-217): for k in {k}, cut 1.35, every
beam Q > k of the reference's list, a checkRecall run; then visit-limit
sweeps (calculate_limits of the first run's average visited count) crossed
with degree-limit sweeps (calculate_limits of the max degree, plus the max
degree); then one "best accuracy" run (k=100, Q=1000, cut=10).  parse_result
(parse_results.h:214-238) keeps, per recall bucket, the run with the best
QPS among runs whose recall lies in [bucket, next bucket]; write_to_csv
(check_nn_recall.h:112-148) writes them in the reference's CSV layout.

checkRecall (check_nn_recall.h:34-110) runs the query set twice (a warm-up,
then the timed pass) and reports QPS = nq / time, recall with ground-truth
ties (recall_with_ties), and stats.h:77-85's average and 99th-percentile
dist_cmps / visited counts.  The search starts from start_point (the fixed
start; the reference's default random start depends on its RNG).
"""
from __future__ import annotations

import csv
import time
from dataclasses import dataclass, field

import numpy as np

from .index import recall_with_ties

BEAMS = [10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40, 45, 50, 55, 60, 65,
         70, 80, 90, 100, 120, 140, 160, 180, 200, 225, 250, 275, 300, 375, 500, 750, 1000]  # check_nn_recall.h:173-175
BUCKETS = [.1, .2, .3, .4, .5, .6, .7, .75, .8, .85, .9, .93, .95, .97, .98, .99, .995, .999, .9995, .9999,
           .99995, .99999]  # check_nn_recall.h:205-207


@dataclass
class NNResult:
    """nn_result (parse_results.h:124-175)."""
    recall: float
    avg_cmps: int
    tail_cmps: int
    avg_visited: int
    tail_visited: int
    QPS: float
    k: int
    beamQ: int
    cut: float
    limit: int
    degree_limit: int
    gtn: int
    num_queries: int

    def line(self):
        return (f"For {self.gtn}@{self.gtn} recall = {self.recall}, QPS = {self.QPS}, Q = {self.beamQ}, "
                f"cut = {self.cut}, visited limit = {self.limit}, degree limit: {self.degree_limit}, "
                f"average visited = {self.avg_visited}, average cmps = {self.avg_cmps}")


@dataclass
class GraphInfo:
    """Graph_ (the CSV's graph header row): name, parameters, size, build time,
    average and max degree."""
    name: str
    params: str
    size: int
    time: float
    avg_deg: float
    max_deg: int


def statistics(s) -> tuple[int, int]:
    """stats.h:77-85: (average, 99th percentile) of a per-query counter, with
    the reference's integer arithmetic: the u32 sum cast to int over a double
    count, truncated; the tail at index (u32)(.99 * (float)n) of the sorted
    values."""
    v = np.sort(np.asarray(s, dtype=np.uint64).astype(np.uint32))
    total = int(v.astype(np.uint64).sum()) & 0xFFFFFFFF  # parlay::reduce over u32 wraps
    total = total - (1 << 32) if total >= (1 << 31) else total  # (int) cast
    avg = int(total / float(len(v))) & 0xFFFFFFFF
    tail_index = int(0.99 * float(np.float32(len(v))))  # double .99 times the float count
    return avg, int(v[tail_index])


def calculate_limits(upper_bound: int) -> list[int]:
    """check_nn_recall.h:150-157: (long)(i * (float)upper_bound * .1) for
    i = 1..9 (a float product, then a double multiply, truncated), duplicates
    removed.  parlay::remove_duplicates (parlay/internal/group_by.h:216) leaves
    the survivors in hash order; they are kept in increasing order here.  The
    order only decides which of two runs with equal QPS parse_result keeps."""
    out = []
    for i in range(1, 10):
        v = int(float(np.float32(i) * np.float32(upper_bound)) * 0.1)
        if v not in out:
            out.append(v)
    return out


def check_recall(idx, Q, gt_ids, gt_d, k, beam, cut=1.35, limit=0, degree_limit=0, start_point=0, qk=None,
                 trav="native", rerank=0) -> NNResult:
    """checkRecall (check_nn_recall.h:34-110) with the fixed start point."""
    qk = k if qk is None else qk
    lim = idx.n if limit == 0 else limit
    dl = idx.R if degree_limit == 0 else degree_limit
    idx.search(Q, qk, beam, cut, limit=lim, degree_limit=dl, starts=(start_point,), trav=trav, rerank=rerank)
    t0 = time.perf_counter()
    ids, _, cmps, vis = idx.search(Q, qk, beam, cut, limit=lim, degree_limit=dl, starts=(start_point,), trav=trav,
                                   rerank=rerank)
    qt = time.perf_counter() - t0
    rec = recall_with_ties(ids[:, :k], gt_ids, gt_d, k)
    ac, tc = statistics(cmps)
    av, tv = statistics(vis)
    # nn_result's k and gtn are both checkRecall's k (check_nn_recall.h:110), also for the k=100 run
    return NNResult(rec, ac, tc, av, tv, Q.shape[0] / qt, k, beam, cut, lim, dl, k, Q.shape[0])


def parse_result(results: list[NNResult], buckets=BUCKETS):
    """parse_results.h:214-238: per bucket b, the best-QPS run among those with
    recall >= b and (unless b is the last bucket or none qualify) <= the
    next bucket."""
    ret, ret_b = [], []
    for i, b in enumerate(buckets):
        cand = [r for r in results if r.recall >= b]
        if i != len(buckets) - 1 and cand:
            cand = [r for r in cand if r.recall <= buckets[i + 1]]
        if cand:
            best = cand[0]
            for r in cand[1:]:
                if best.QPS < r.QPS:  # parlay::max_element keeps the first maximum
                    best = r
            ret.append(best)
            ret_b.append(b)
    return ret, ret_b


def write_to_csv(path, buckets, results: list[NNResult], G: GraphInfo):
    """check_nn_recall.h:112-148 (csvfile.h separator ",")."""
    with open(path, "a", newline="") as f:
        w = csv.writer(f)
        w.writerow(["GRAPH", "Parameters", "Size", "Build time", "Avg degree", "Max degree"])
        w.writerow([G.name, G.params, G.size, G.time, G.avg_deg, G.max_deg])
        w.writerow([])
        w.writerow(["Num queries", "Target recall", "Actual recall", "QPS", "Average Cmps", "Tail Cmps",
                    "Average Visited", "Tail Visited", "k", "Q", "cut"])
        for b, r in zip(buckets, results):
            w.writerow([r.num_queries, b, r.recall, r.QPS, r.avg_cmps, r.tail_cmps, r.avg_visited, r.tail_visited,
                        r.k, r.beamQ, r.cut])
        w.writerow([])
        w.writerow([])


@dataclass
class SweepReport:
    results: list = field(default_factory=list)   # every run, in the reference's order
    best: list = field(default_factory=list)      # parse_result's picks
    buckets: list = field(default_factory=list)


def search_and_parse(idx, Q, gt_ids, gt_d, k=10, start_point=0, res_file=None, graph: GraphInfo | None = None,
                     beams=BEAMS, log=print) -> SweepReport:
    """check_nn_recall.h:159-217."""
    rep = SweepReport()
    for Qb in beams:
        if Qb > k:
            rep.results.append(check_recall(idx, Q, gt_ids, gt_d, k, Qb, 1.35, start_point=start_point))
    limits = calculate_limits(rep.results[0].avg_visited)
    degree_limits = calculate_limits(idx.R) + [idx.R]
    # A zero limit (a first run visiting < 10 vertices on average, or R < 10)
    # stops the reference's search before its first expansion; the C-ABI reads
    # 0 as "no limit", so such a limit runs as 1 here.
    for lim in limits:
        for dl in degree_limits:
            rep.results.append(check_recall(idx, Q, gt_ids, gt_d, k, max(lim, k), 1.35, limit=max(lim, 1),
                                            degree_limit=max(dl, 1), start_point=start_point))
    # "best accuracy": QueryParams(100, 1000, 10.0, n, max_degree), recall@k
    rep.results.append(check_recall(idx, Q, gt_ids, gt_d, k, 1000, 10.0, start_point=start_point, qk=100))
    rep.best, rep.buckets = parse_result(rep.results)
    if log:
        for r in rep.best:
            log(r.line())
    if res_file is not None and graph is not None:
        write_to_csv(res_file, rep.buckets, rep.best, graph)
    return rep
