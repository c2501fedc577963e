"""Host-side mirror of the reference's build/search interface over libddg.

``Index`` is the thin owner of one device handle (vectors + graph).  The
reference-named classes and functions below (``PointRange``, ``Graph``,
``BuildParams``, ``QueryParams``, ``stats``, ``knn_index``, ``searchAll``,
``checkRecall``) keep the argument meaning of include/index.h,
include/utils/beamSearch.h and include/utils/check_nn_recall.h; where the
reference prints and aborts, these raise ``DdgError``.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import formats
from ._lib import (DTYPES, METRICS, TRAV, BuildParamsC, DdgError, QueryParamsC, check, lib)

_U32 = np.uint32


def _c(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


class Index:
    """One device copy of a PointRange (aligned rows) and of its Graph."""

    def __init__(self, rows: np.ndarray, d: int, metric: str = "l2", R: int = 32):
        rows = np.ascontiguousarray(rows)
        if rows.ndim != 2 or rows.dtype.name not in DTYPES:
            raise ValueError("rows must be a 2-D uint8/int8/float32 array")
        self.rows = rows
        self.d = int(d)
        self.ad = rows.shape[1]
        self.n = rows.shape[0]
        self.R = int(R)
        self.dtype = rows.dtype.name
        self.metric = metric
        h = C.c_void_p()
        check(lib().ddg_create(DTYPES[self.dtype], METRICS[metric], self.n, self.d, self.ad, _c(rows), self.R,
                               C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().ddg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st):
        check(st, self.h)

    # ------------------------------------------------------------- graph
    def get_graph(self) -> np.ndarray:
        out = np.empty((self.n, self.R + 1), _U32)
        self._chk(lib().ddg_graph_get(self.h, _c(out)))
        return out

    def set_graph(self, rows: np.ndarray):
        rows = np.ascontiguousarray(rows, dtype=_U32)
        if rows.shape != (self.n, self.R + 1):
            raise ValueError(f"graph rows must be {(self.n, self.R + 1)}")
        self._chk(lib().ddg_graph_set(self.h, _c(rows)))

    def get_rows(self, ids) -> np.ndarray:
        """Graph rows of `ids` ([degree, nbrs...] each)."""
        ids = np.ascontiguousarray(np.asarray(ids, _U32))
        out = np.empty((len(ids), self.R + 1), _U32)
        self._chk(lib().ddg_graph_get_rows(self.h, _c(ids), len(ids), _c(out)))
        return out

    def set_rows(self, ids, rows: np.ndarray):
        """Replaces the graph rows of `ids` (G[i].update_neighbors, batched)."""
        ids = np.ascontiguousarray(np.asarray(ids, _U32))
        rows = np.ascontiguousarray(rows, dtype=_U32)
        if rows.shape != (len(ids), self.R + 1):
            raise ValueError(f"rows must be {(len(ids), self.R + 1)}")
        self._chk(lib().ddg_graph_set_rows(self.h, _c(ids), len(ids), _c(rows)))

    # ------------------------------------------------------------- build
    def build(self, R=None, L=64, alpha=1.2, two_pass=False, start_point=0, base=2.0, max_fraction=1.0,
              max_build=0, final_sort=True, bsp=False, visited_counts=False):
        R = self.R if R is None else R
        bp = BuildParamsC(R, L, alpha, int(two_pass), start_point, base, max_fraction, max_build, int(final_sort),
                          int(bsp))
        vc = np.zeros(self.n, _U32) if visited_counts else None
        secs = C.c_double(0)
        self._chk(lib().ddg_build(self.h, C.byref(bp), _c(vc), C.byref(secs)))
        return (secs.value, vc) if visited_counts else secs.value

    def batch_insert(self, inserts, alpha, L, start_point=0, base=2.0, max_fraction=0.02, random_order=False):
        ins = np.ascontiguousarray(np.asarray(inserts, _U32))
        self._chk(lib().ddg_batch_insert(self.h, _c(ins), len(ins), alpha, L, start_point, int(bool(random_order)),
                                         base, max_fraction))

    def sort_rows(self):
        self._chk(lib().ddg_sort_rows(self.h))

    def robust_prune(self, pts, cand_lists, cand_dists=None, alpha=1.2, add=True, R=None):
        R = self.R if R is None else R
        pts = np.ascontiguousarray(np.asarray(pts, _U32))
        off = np.zeros(len(pts) + 1, np.uint64)
        off[1:] = np.cumsum([len(c) for c in cand_lists])
        ids = np.ascontiguousarray(np.concatenate([np.asarray(c, _U32) for c in cand_lists]) if len(pts) else
                                   np.zeros(0, _U32))
        dd = None
        if cand_dists is not None:
            dd = np.ascontiguousarray(np.concatenate([np.asarray(c, np.float32) for c in cand_dists]))
        out = np.zeros((max(len(pts), 1), R), _U32)
        ln = np.zeros(max(len(pts), 1), _U32)
        self._chk(lib().ddg_robust_prune(self.h, len(pts), _c(pts), _c(off), _c(ids), _c(dd), alpha, int(add), R,
                                         _c(out), _c(ln)))
        return [out[i, :ln[i]].copy() for i in range(len(pts))]

    # ------------------------------------------------------------ search
    def search(self, Q: np.ndarray, k=10, beam=50, cut=1.35, limit=0, degree_limit=0, starts=(0,),
               per_query_beam=None, query_ids=None, trav="native", rerank=0):
        Q = np.ascontiguousarray(Q)
        if Q.dtype != self.rows.dtype or Q.shape[1] != self.ad:
            raise ValueError("queries must match the base dtype and aligned dimension")
        nq = Q.shape[0]
        qp = QueryParamsC(k, beam, cut, limit, degree_limit, TRAV[trav], rerank)
        st = np.ascontiguousarray(np.asarray(starts, _U32))
        pqb = None if per_query_beam is None else np.ascontiguousarray(np.asarray(per_query_beam, _U32))
        qid = None if query_ids is None else np.ascontiguousarray(np.asarray(query_ids, _U32))
        ids = np.empty((nq, k), _U32)
        dists = np.empty((nq, k), np.float32)
        cmps = np.empty(nq, np.uint64)
        vis = np.empty(nq, _U32)
        self._chk(lib().ddg_search(self.h, _c(Q), nq, _c(qid), _c(st), len(st), C.byref(qp), _c(pqb), _c(ids),
                                   _c(dists), _c(cmps), _c(vis)))
        return ids, dists, cmps, vis

    def beam_search(self, Q: np.ndarray, k=0, beam=50, cut=1.35, limit=0, degree_limit=0, starts=(0,),
                    query_ids=None, vcap=None):
        """beam_search (beamSearch.h:41-184) of every row of Q: per query
        ((frontier ids, frontier dists), (visited ids, visited dists),
        dist_cmps), both lists sorted by (distance, id)."""
        Q = np.ascontiguousarray(Q)
        if Q.dtype != self.rows.dtype or Q.shape[1] != self.ad:
            raise ValueError("queries must match the base dtype and aligned dimension")
        nq = Q.shape[0]
        st = np.ascontiguousarray(np.asarray(starts, _U32))
        qid = None if query_ids is None else np.ascontiguousarray(np.asarray(query_ids, _U32))
        fcap = max(beam, len(st))
        vcap = 2 * beam + 64 if vcap is None else vcap
        qp = QueryParamsC(k, beam, cut, limit, degree_limit, TRAV["native"], 0)
        while True:
            fi = np.zeros((max(nq, 1), fcap), _U32)
            fd = np.zeros((max(nq, 1), fcap), np.float32)
            fl = np.zeros(max(nq, 1), _U32)
            vi = np.zeros((max(nq, 1), vcap), _U32)
            vd = np.zeros((max(nq, 1), vcap), np.float32)
            vl = np.zeros(max(nq, 1), _U32)
            cm = np.zeros(max(nq, 1), np.uint64)
            st_ = lib().ddg_beam_search(self.h, _c(Q), nq, _c(qid), _c(st), len(st), C.byref(qp), fcap, _c(fi),
                                        _c(fd), _c(fl), vcap, _c(vi), _c(vd), _c(vl), _c(cm))
            if st_ != 0 and nq and int(vl[:nq].max()) > vcap:
                vcap = int(vl[:nq].max())  # the visited lists were longer: retry with room for them
                continue
            self._chk(st_)
            break
        return [((fi[i, :fl[i]].copy(), fd[i, :fl[i]].copy()), (vi[i, :vl[i]].copy(), vd[i, :vl[i]].copy()),
                 int(cm[i])) for i in range(nq)]

    def search_device(self, Q, k, beam, out_ids, out_dists=None, cmps=None, visited=None, cut=1.35, limit=0,
                      degree_limit=0, starts=None, per_query_beam=None, beam_max=None, query_ids=None,
                      trav="native", rerank=0):
        """Search over torch CUDA tensors (device-resident inputs and outputs)."""
        if starts is None:
            raise ValueError("search_device needs a device tensor of start points")
        want = {"uint8": "torch.uint8", "int8": "torch.int8", "float32": "torch.float32"}[self.dtype]
        if str(Q.dtype) != want or Q.dim() != 2 or Q.shape[1] != self.ad or not Q.is_contiguous():
            raise ValueError(f"queries must be a contiguous (nq, {self.ad}) {want} tensor")
        nq = Q.shape[0]
        for name, t, size in (("out_ids", out_ids, nq * k), ("out_dists", out_dists, nq * k), ("cmps", cmps, nq),
                              ("visited", visited, nq), ("per_query_beam", per_query_beam, nq),
                              ("query_ids", query_ids, nq)):
            if t is not None and (t.numel() < size or not t.is_contiguous() or t.element_size() not in (4, 8)):
                raise ValueError(f"{name} must be a contiguous tensor of at least {size} elements")
        if cmps is not None and cmps.element_size() != 8:
            raise ValueError("cmps must hold 64-bit counters")
        qp = QueryParamsC(k, beam, cut, limit, degree_limit, TRAV[trav], rerank)
        p = lambda t: None if t is None else C.c_void_p(t.data_ptr())
        beam_max = beam if beam_max is None else beam_max
        nst = starts.numel()
        self._chk(lib().ddg_search_device(self.h, p(Q), Q.shape[0], p(query_ids), p(starts), nst, C.byref(qp),
                                          p(per_query_beam), beam_max, p(out_ids), p(out_dists), p(cmps),
                                          p(visited)))

    def autotune_search(self, Q: np.ndarray, k=10, beam=50, cut=1.35, trav="native", rerank=0):
        """Times the result-preserving search launch shapes on Q and keeps the
        fastest for later searches with this traversal mode and beam class."""
        Q = np.ascontiguousarray(Q)
        if Q.dtype != self.rows.dtype or Q.shape[1] != self.ad:
            raise ValueError("queries must match the base dtype and aligned dimension")
        qp = QueryParamsC(k, beam, cut, 0, 0, TRAV[trav], rerank)
        out = np.zeros(4, _U32)
        ms = np.zeros(2, np.float64)
        self._chk(lib().ddg_autotune_search(self.h, _c(Q), Q.shape[0], C.byref(qp), _c(out), _c(ms)))
        return {"global_filter": bool(out[0] & 1), "staged_rows": int(out[0] >> 8), "waves_per_block": int(out[1]),
                "blocks_per_cu": int(out[2]), "tried": int(out[3]), "default_ms": float(ms[0]),
                "tuned_ms": float(ms[1])}

    def search_shape(self, beam, trav="native"):
        """The launch shape a search with this beam would use now: dict of
        global_filter, staged_rows, waves_per_block, blocks_per_cu and source
        ("default", "handle" or "process": where a tuned shape came from)."""
        qp = QueryParamsC(0, beam, 1.35, 0, 0, TRAV[trav], 0)
        out = np.zeros(4, _U32)
        self._chk(lib().ddg_search_shape(self.h, C.byref(qp), _c(out)))
        return {"global_filter": bool(out[0] & 1), "staged_rows": int(out[0] >> 8), "waves_per_block": int(out[1]),
                "blocks_per_cu": int(out[2]), "source": ["default", "handle", "process"][int(out[3])]}

    def prepare_traversal(self, trav):
        """Builds the BF16 / FP8 traversal copy; returns its device seconds."""
        sec = C.c_double(0.0)
        self._chk(lib().ddg_prepare_traversal(self.h, TRAV[trav], C.byref(sec)))
        return sec.value

    def distances(self, Q: np.ndarray, ids, trav="native"):
        """Distances of query i to ids[i, j] in the arithmetic of `trav`."""
        Q = np.ascontiguousarray(Q)
        if Q.dtype != self.rows.dtype or Q.shape[1] != self.ad:
            raise ValueError("queries must match the base dtype and aligned dimension")
        ids = np.ascontiguousarray(np.asarray(ids, _U32))
        if ids.ndim != 2 or ids.shape[0] != Q.shape[0]:
            raise ValueError("ids must be (nq, m)")
        out = np.empty(ids.shape, np.float32)
        self._chk(lib().ddg_distances(self.h, TRAV[trav], _c(Q), Q.shape[0], _c(ids), ids.shape[1], _c(out)))
        return out

    def insert_search(self, ids, L, start_point=0, vcap=None):
        ids = np.ascontiguousarray(np.asarray(ids, _U32))
        vcap = 4 * L + 64 if vcap is None else vcap
        m = len(ids)
        vi = np.zeros((max(m, 1), vcap), _U32)
        vd = np.zeros((max(m, 1), vcap), np.float32)
        vl = np.zeros(max(m, 1), _U32)
        cm = np.zeros(max(m, 1), np.uint64)
        self._chk(lib().ddg_insert_search(self.h, _c(ids), m, L, start_point, vcap, _c(vi), _c(vd), _c(vl), _c(cm)))
        return [(vi[i, :vl[i]].copy(), vd[i, :vl[i]].copy()) for i in range(m)], cm[:m].copy()

    def ground_truth(self, Q: np.ndarray, k: int):
        Q = np.ascontiguousarray(Q)
        if Q.dtype != self.rows.dtype or Q.shape[1] != self.ad:
            raise ValueError("queries must match the base dtype and aligned dimension")
        ids = np.empty((Q.shape[0], k), _U32)
        d = np.empty((Q.shape[0], k), np.float64)
        self._chk(lib().ddg_ground_truth(self.h, _c(Q), Q.shape[0], k, _c(ids), _c(d)))
        return ids, d

    def profile(self):
        out = np.zeros(9, np.float64)
        self._chk(lib().ddg_last_profile(self.h, _c(out)))
        keys = ["search_ms", "prune_ms", "edges_ms", "reverse_ms", "sort_ms", "dense_ms", "search_bytes",
                "dist_cmps", "visited"]
        return dict(zip(keys, out.tolist()))


# ====================================================================
# Mirrors of the reference's types and entry points
# ====================================================================
@dataclass
class BuildParams:
    """types.h:103-142 (Vamana fields)."""
    R: int
    L: int
    alpha: float
    two_pass: bool = False


@dataclass
class QueryParams:
    """types.h:145-156."""
    k: int
    beamSize: int
    cut: float
    limit: int
    degree_limit: int


class PointRange:
    """point_range.h:53-113: aligned rows plus the Point type's metric."""

    def __init__(self, rows_or_file, dtype: str | None = None, metric: str = "l2", d: int | None = None):
        if isinstance(rows_or_file, str):
            self.values, self.dims = formats.read_vectors(rows_or_file, dtype)
        else:
            X = np.asarray(rows_or_file)
            self.dims = X.shape[1] if d is None else d
            self.values = formats.align_rows(X) if d is None else np.ascontiguousarray(X)
        self.metric = metric
        self._index = None

    def size(self):
        return self.values.shape[0]

    def dimension(self):
        return self.dims

    def aligned_dimension(self):
        return self.values.shape[1]


class Graph:
    """graph.h:119-209: n x (R+1) rows [deg, nbrs...]."""

    def __init__(self, maxDeg_or_file, n: int | None = None):
        if isinstance(maxDeg_or_file, str):
            self.rows = formats.read_graph(maxDeg_or_file)
        else:
            self.rows = np.zeros((n, maxDeg_or_file + 1), _U32)

    def size(self):
        return self.rows.shape[0]

    def max_degree(self):
        return self.rows.shape[1] - 1

    def __getitem__(self, i):
        return self.rows[i, 1:1 + self.rows[i, 0]]

    def save(self, path: str):
        formats.write_graph(path, self.rows)


@dataclass
class stats:
    """stats.h:53-88: per-point visited / distance counters."""
    visited: np.ndarray = field(default_factory=lambda: np.zeros(0, _U32))
    distances: np.ndarray = field(default_factory=lambda: np.zeros(0, _U32))

    @classmethod
    def of_size(cls, n):
        return cls(np.zeros(n, _U32), np.zeros(n, _U32))

    @staticmethod
    def statistics(s):
        s = np.sort(np.asarray(s))
        avg = int(int(s.sum()) / float(len(s))) if len(s) else 0
        tail = int(s[int(0.99 * float(len(s)))]) if len(s) else 0
        return avg, tail

    def visited_stats(self):
        return self.statistics(self.visited)

    def dist_stats(self):
        return self.statistics(self.distances)


def _device_index(Points: PointRange, G: Graph) -> Index:
    idx = Points._index
    if idx is None or idx.R != G.max_degree():
        if idx is not None:
            idx.close()
        idx = Index(Points.values, Points.dims, Points.metric, G.max_degree())
        Points._index = idx
    return idx


class knn_index:
    """index.h:43-369 (Vamana)."""

    def __init__(self, BP: BuildParams):
        self.BP = BP
        self.start_point = 0
        self.delete_set = set()  # index.h:52; nothing on the build or search path reads it

    def get_start(self):
        return self.start_point

    def lazy_delete(self, deletes, G: Graph):
        """index.h:139-164: marks points deleted.  An id above G.size() is an
        error (the reference aborts; its bound lets G.size() through); the
        start point is refused with a message."""
        for p in ([deletes] if np.isscalar(deletes) else deletes):
            p = int(p)
            if p > G.size():
                raise DdgError(1, f"ERROR: invalid point {p} given to lazy_delete")
            if p == self.start_point:
                print("Deleting start_point not permitted; continuing")
                continue
            self.delete_set.add(p)

    def build_index(self, G: Graph, Points: PointRange, BuildStats: stats | None = None):
        """index.h:125-137: start point 0, all points, final per-row sort."""
        self.start_point = 0
        idx = _device_index(Points, G)
        idx.set_graph(G.rows)
        secs, vc = idx.build(R=self.BP.R, L=self.BP.L, alpha=self.BP.alpha, two_pass=self.BP.two_pass,
                             visited_counts=True)
        G.rows[...] = idx.get_graph()
        if BuildStats is not None and len(BuildStats.visited) == len(vc):
            BuildStats.visited += vc
        return secs

    def batch_insert(self, inserts, G: Graph, Points: PointRange, BuildStats=None, alpha=1.2, random_order=False,
                     base=2.0, max_fraction=0.02, print_=True):
        """index.h:213-341 (random_order: parlay::random_permutation's order)."""
        idx = _device_index(Points, G)
        idx.set_graph(G.rows)
        idx.batch_insert(inserts, alpha, self.BP.L, self.start_point, base, max_fraction, random_order)
        G.rows[...] = idx.get_graph()

    def robustPrune(self, p, cand, G: Graph, Points: PointRange, alpha, add=True):
        """index.h:63-123: ``cand`` is a list of ids or of (id, distance) pairs."""
        idx = _device_index(Points, G)
        idx.set_graph(G.rows)
        cand = list(cand)
        if cand and isinstance(cand[0], (tuple, list)):
            ids = [c[0] for c in cand]
            ds = [c[1] for c in cand]
            return idx.robust_prune([p], [ids], [ds], alpha, add, self.BP.R)[0]
        return idx.robust_prune([p], [cand], None, alpha, add, self.BP.R)[0]


def searchAll(Query_Points: PointRange, G: Graph, Base_Points: PointRange, QueryStats: stats | None,
              starting_point, QP: QueryParams, per_query_beam=None):
    """beamSearch.h:314-346 (query i has p.id() = i); returns nq x k ids."""
    idx = _device_index(Base_Points, G)
    idx.set_graph(G.rows)
    starts = [starting_point] if np.isscalar(starting_point) else list(starting_point)
    ids, _, cmps, vis = idx.search(Query_Points.values, QP.k, QP.beamSize, QP.cut, QP.limit, QP.degree_limit,
                                   starts, per_query_beam)
    if QueryStats is not None and len(QueryStats.visited) == len(vis):
        QueryStats.visited += vis
        QueryStats.distances += cmps.astype(_U32)
    return ids


def beam_search(p, G: Graph, Points: PointRange, starting_point, QP: QueryParams, p_id=None):
    """beamSearch.h:41-184: ((frontier, visited), dist_cmps) for one point p
    (an aligned row of Points' dtype); p_id is p.id() for the self-skip
    (None: no point of Points is skipped).  Lists are (ids, distances)."""
    idx = _device_index(Points, G)
    idx.set_graph(G.rows)
    starts = [starting_point] if np.isscalar(starting_point) else list(starting_point)
    row = np.ascontiguousarray(np.asarray(p, Points.values.dtype).reshape(1, -1))
    qid = None if p_id is None else [p_id]
    (fr, vis, cmps), = idx.beam_search(row, QP.k, QP.beamSize, QP.cut, QP.limit, QP.degree_limit, starts,
                                       query_ids=qid if qid is not None else [0xFFFFFFFF])
    return (fr, vis), cmps


def quantized_mips_distances(codes: np.ndarray, queries: np.ndarray, max_coord: float, min_coord: float):
    """Quantized_Mips_Point<T>::distance (mips_point.h:111-187) of every
    (query, code row) pair on the GPU, nq x n: float32 queries are
    Mips_Point<float> queries, uint8 / int8 ones are quantized queries."""
    codes = np.ascontiguousarray(codes)
    queries = np.ascontiguousarray(queries)
    if codes.dtype.name not in ("uint8", "int8") or codes.ndim != 2:
        raise ValueError("codes must be a 2-D uint8 or int8 array")
    qq = queries.dtype != np.float32
    if (qq and queries.dtype != codes.dtype) or queries.ndim != 2 or queries.shape[1] < codes.shape[1]:
        raise ValueError("queries must be float32 or of the codes' dtype, with at least d columns")
    out = np.empty((queries.shape[0], codes.shape[0]), np.float32)
    check(lib().ddg_quantized_mips_distances(DTYPES[codes.dtype.name], codes.shape[1], max_coord, min_coord,
                                             _c(queries), int(qq), queries.shape[0], queries.shape[1], _c(codes),
                                             codes.shape[0], codes.shape[1], _c(out)))
    return out


def recall_with_ties(found: np.ndarray, gt_ids: np.ndarray, gt_dists: np.ndarray, k: int) -> float:
    """check_nn_recall.h:86-109: the first k ground-truth ids plus every later
    one at the k-th distance count as correct answers."""
    n = found.shape[0]
    gt_ids = np.asarray(gt_ids)
    gt_dists = np.asarray(gt_dists)
    # ground-truth entry l counts when l < k or its distance equals the k-th
    counted = np.ones(gt_ids.shape, bool)
    counted[:, k:] = gt_dists[:, k:] == gt_dists[:, k - 1:k]
    rep = np.asarray(found)[:, :k].astype(np.int64)
    hit = (gt_ids.astype(np.int64)[:, :, None] == rep[:, None, :]).any(axis=2)
    return int((hit & counted).sum()) / float(k * n)


@dataclass
class nn_result:
    recall: float
    QPS: float
    avg_cmps: float
    tail_cmps: float
    avg_visited: float
    tail_visited: float
    k: int
    beamQ: int
    cut: float
    num_queries: int


def checkRecall(G: Graph, Base_Points: PointRange, Query_Points: PointRange, GT, random: bool, start_point, k,
                QP: QueryParams):
    """check_nn_recall.h:34-115: a warm-up search, a timed search, recall with
    ground-truth ties and QPS = nq / t."""
    if random:
        raise DdgError(1, "random start points are not supported; use a fixed start point")
    gt_ids, gt_d = GT
    if k > gt_ids.shape[1]:
        raise DdgError(1, f"{k}@{k} too large for ground truth data of size {gt_ids.shape[1]}")
    searchAll(Query_Points, G, Base_Points, None, start_point, QP)
    qs = stats.of_size(Query_Points.size())
    t = time.perf_counter()
    ids = searchAll(Query_Points, G, Base_Points, qs, start_point, QP)
    dt = time.perf_counter() - t
    rec = recall_with_ties(ids, gt_ids, gt_d, k)
    ac, tc = qs.dist_stats()
    av, tv = qs.visited_stats()
    return nn_result(rec, Query_Points.size() / dt, ac, tc, av, tv, k, QP.beamSize, QP.cut, Query_Points.size())
