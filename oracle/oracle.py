"""ORACLE — test infrastructure only.

ctypes wrapper over oracle/liboracle.so, the CPU restatement of the
reference's build and search path (see the header of oracle/oracle.cpp for
the function-by-function anchors and what pins each one).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
and only as the checker / the timed CPU baseline.  The product package never
imports it.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

DTYPES = {"uint8": 0, "int8": 1, "float32": 2}
NP_DTYPES = {"uint8": np.uint8, "int8": np.int8, "float32": np.float32}
METRICS = {"l2": 0, "mips": 1}
TIE_REFERENCE, TIE_ID = 0, 1


class _Points(C.Structure):
    _fields_ = [("dtype", C.c_int), ("metric", C.c_int), ("n", C.c_uint64), ("d", C.c_uint32),
                ("ad", C.c_uint32), ("rows", C.c_void_p)]


class _Graph(C.Structure):
    _fields_ = [("n", C.c_uint64), ("R", C.c_uint32), ("rows", C.c_void_p)]


class _BuildParams(C.Structure):
    _fields_ = [("R", C.c_uint32), ("L", C.c_uint32), ("alpha", C.c_double), ("two_pass", C.c_int),
                ("start_point", C.c_uint32), ("base", C.c_double), ("max_fraction", C.c_double),
                ("max_build", C.c_uint64), ("final_sort", C.c_int), ("bsp", C.c_int)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load(LIB_PATH)
    return _lib


@contextlib.contextmanager
def using(path):
    """Runs the wrapper's functions against another build of oracle.cpp (the
    mutation tests compile deliberately broken copies)."""
    global _lib
    saved = lib()
    _lib = load(path)
    try:
        yield
    finally:
        _lib = saved


def load(path):
    """A ctypes handle on a build of oracle.cpp with every signature set."""
    L = C.CDLL(path)
    P = C.POINTER
    L.orc_hash64_2.restype = C.c_uint64
    L.orc_hash64_2.argtypes = [C.c_uint64]
    L.orc_random_permutation.argtypes = [C.c_uint64, C.c_void_p]
    L.orc_random_permutation.restype = None
    L.orc_beam_search.argtypes = [P(_Points), P(_Graph), C.c_void_p, C.c_int64, C.c_void_p, C.c_uint32,
                                     C.c_int64, C.c_int64, C.c_double, C.c_int64, C.c_int64,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
    L.orc_search_all.argtypes = [P(_Points), P(_Graph), C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint32,
                                    C.c_int64, C.c_int64, C.c_double, C.c_int64, C.c_int64, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    L.orc_robust_prune.argtypes = [P(_Points), P(_Graph), C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32,
                                      C.c_double, C.c_int, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p]
    L.orc_schedule.argtypes = [C.c_uint64, C.c_double, C.c_double, C.c_uint64, C.c_void_p, C.c_void_p,
                                  C.c_uint64, C.c_void_p]
    L.orc_group_by_key.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p]
    L.orc_batch_insert.argtypes = [P(_Points), P(_Graph), C.c_void_p, C.c_uint64, C.c_double, C.c_double,
                                      C.c_double, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_int,
                                      C.c_void_p, C.c_void_p]
    L.orc_build.argtypes = [P(_Points), P(_Graph), P(_BuildParams), C.c_int, C.c_int, C.c_void_p,
                               C.c_void_p]
    L.orc_sort_rows.argtypes = [P(_Points), P(_Graph), C.c_int, C.c_int]
    L.orc_ground_truth.argtypes = [P(_Points), C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p, C.c_void_p,
                                      C.c_int]
    L.orc_distances.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32,
                                   C.c_uint32, C.c_void_p]
    L.orc_quantized_mips.argtypes = [C.c_int, C.c_uint32, C.c_float, C.c_float, C.c_void_p, C.c_int, C.c_uint64,
                                     C.c_void_p, C.c_uint64, C.c_void_p]
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data


class Points:
    """Aligned rows (point_range.h:45-50, 80) plus dtype/metric."""

    def __init__(self, rows: np.ndarray, d: int, metric: str = "l2"):
        assert rows.ndim == 2 and rows.flags.c_contiguous
        self.rows = rows
        self.d = d
        self.dtype = rows.dtype.name
        self.metric = metric
        self.s = _Points(DTYPES[self.dtype], METRICS[metric], rows.shape[0], d, rows.shape[1], rows.ctypes.data)

    @property
    def n(self):
        return self.rows.shape[0]


class Graph:
    def __init__(self, rows: np.ndarray):
        assert rows.dtype == np.uint32 and rows.ndim == 2 and rows.flags.c_contiguous
        self.rows = rows
        self.s = _Graph(rows.shape[0], rows.shape[1] - 1, rows.ctypes.data)

    @property
    def R(self):
        return self.rows.shape[1] - 1


def hash64_2(x: int) -> int:
    return int(lib().orc_hash64_2(x))


def random_permutation(n: int) -> np.ndarray:
    """parlay::random_permutation<int>(n) as batch_insert's random_order draws it."""
    out = np.empty(n, np.int32)
    lib().orc_random_permutation(n, _ptr(out) if n else None)
    return out


def distances(dtype, metric, A: np.ndarray, B: np.ndarray, d: int) -> np.ndarray:
    A = np.ascontiguousarray(A)
    B = np.ascontiguousarray(B)
    out = np.empty(A.shape[0], np.float32)
    lib().orc_distances(DTYPES[dtype], METRICS[metric], A.ctypes.data, B.ctypes.data, A.shape[0], d,
                        A.shape[1], out.ctypes.data)
    return out


def beam_search(P: Points, G: Graph, query: np.ndarray, self_id: int, starts, k, beam, cut, limit, degree_limit):
    starts = np.ascontiguousarray(np.asarray(starts, np.uint32))
    query = np.ascontiguousarray(query)
    cap_f, cap_v = int(beam) + 8, 1 << 16
    fi = np.empty(cap_f, np.uint32); fd = np.empty(cap_f, np.float32)
    vi = np.empty(cap_v, np.uint32); vd = np.empty(cap_v, np.float32)
    fl = np.zeros(1, np.uint32); vl = np.zeros(1, np.uint32); cm = np.zeros(1, np.uint64)
    r = lib().orc_beam_search(C.byref(P.s), C.byref(G.s), query.ctypes.data, self_id, starts.ctypes.data,
                              len(starts), k, beam, cut, limit, degree_limit, fi.ctypes.data, fd.ctypes.data,
                              fl.ctypes.data, cap_f, vi.ctypes.data, vd.ctypes.data, vl.ctypes.data, cap_v,
                              cm.ctypes.data)
    assert r == 0
    return (fi[:fl[0]].copy(), fd[:fl[0]].copy()), (vi[:vl[0]].copy(), vd[:vl[0]].copy()), int(cm[0])


def search_all(P: Points, G: Graph, Q: np.ndarray, k, beam, cut=1.35, limit=None, degree_limit=None, starts=(0,),
               per_query_beam=None, nthreads=0):
    Q = np.ascontiguousarray(Q)
    assert Q.shape[1] == P.rows.shape[1] and Q.dtype == P.rows.dtype
    nq = Q.shape[0]
    starts = np.ascontiguousarray(np.asarray(starts, np.uint32))
    limit = P.n if limit is None else limit
    degree_limit = G.R if degree_limit is None else degree_limit
    pqb = None if per_query_beam is None else np.ascontiguousarray(np.asarray(per_query_beam, np.uint32))
    ids = np.empty((nq, k), np.uint32); dists = np.empty((nq, k), np.float32)
    cmps = np.empty(nq, np.uint64); vis = np.empty(nq, np.uint32)
    r = lib().orc_search_all(C.byref(P.s), C.byref(G.s), Q.ctypes.data, nq, starts.ctypes.data, len(starts), k,
                             beam, cut, limit, degree_limit, _ptr(pqb), ids.ctypes.data, dists.ctypes.data,
                             cmps.ctypes.data, vis.ctypes.data, nthreads)
    if r != 0:
        raise ValueError("search_all: k > beam")
    return ids, dists, cmps, vis


def robust_prune(P: Points, G: Graph, p: int, cand_ids, cand_dists, alpha, add=True, R=None, tie_mode=TIE_ID):
    cand_ids = np.ascontiguousarray(np.asarray(cand_ids, np.uint32))
    cd = None if cand_dists is None else np.ascontiguousarray(np.asarray(cand_dists, np.float32))
    R = G.R if R is None else R
    out = np.empty(max(R, 1), np.uint32); ol = np.zeros(1, np.uint32)
    lib().orc_robust_prune(C.byref(P.s), C.byref(G.s), p, cand_ids.ctypes.data, _ptr(cd), len(cand_ids), alpha,
                           int(add), R, tie_mode, out.ctypes.data, ol.ctypes.data)
    return out[:ol[0]].copy()


def schedule(m, base=2.0, max_fraction=1.0, max_batch_base=None):
    max_batch_base = m if max_batch_base is None else max_batch_base
    cap = 4096
    f = np.empty(cap, np.uint64); c = np.empty(cap, np.uint64); nb = np.zeros(1, np.uint64)
    assert lib().orc_schedule(m, base, max_fraction, max_batch_base, f.ctypes.data, c.ctypes.data, cap,
                              nb.ctypes.data) == 0
    return list(zip(f[:nb[0]].tolist(), c[:nb[0]].tolist()))


def group_by_key(keys, vals):
    keys = np.ascontiguousarray(np.asarray(keys, np.uint32))
    vals = np.ascontiguousarray(np.asarray(vals, np.uint32))
    n = len(keys)
    ok = np.empty(max(n, 1), np.uint32); oo = np.empty(n + 1, np.uint64); ov = np.empty(max(n, 1), np.uint32)
    ng = np.zeros(1, np.uint64)
    lib().orc_group_by_key(keys.ctypes.data, vals.ctypes.data, n, ok.ctypes.data, oo.ctypes.data, ov.ctypes.data,
                           ng.ctypes.data)
    g = int(ng[0])
    return {int(ok[i]): ov[oo[i]:oo[i + 1]].tolist() for i in range(g)}


def build(P: Points, R, L, alpha, two_pass=False, start_point=0, base=2.0, max_fraction=1.0, max_build=0,
          final_sort=True, bsp=False, tie_mode=TIE_REFERENCE, nthreads=0, return_stats=False):
    rows = np.zeros((P.n, R + 1), np.uint32)
    G = Graph(rows)
    bp = _BuildParams(R, L, alpha, int(two_pass), start_point, base, max_fraction, max_build, int(final_sort),
                      int(bsp))
    vis = np.zeros(P.n, np.uint32)
    st = np.zeros(6, np.uint64)
    r = lib().orc_build(C.byref(P.s), C.byref(G.s), C.byref(bp), tie_mode, nthreads, vis.ctypes.data,
                        st.ctypes.data)
    if r != 0:
        raise RuntimeError(f"oracle build failed: {r}")
    if return_stats:
        return G, vis, dict(zip(["dist_cmps", "visited", "prunes_fwd", "prunes_rev", "appends", "pairs"],
                                st.tolist()))
    return G


def batch_insert(P: Points, G: Graph, inserts, alpha, L, R=None, base=2.0, max_fraction=0.02, start=0,
                 tie_mode=TIE_REFERENCE, nthreads=0):
    ins = np.ascontiguousarray(np.asarray(inserts, np.uint32))
    R = G.R if R is None else R
    r = lib().orc_batch_insert(C.byref(P.s), C.byref(G.s), ins.ctypes.data, len(ins), alpha, base, max_fraction,
                               start, L, R, tie_mode, nthreads, None, None)
    if r != 0:
        raise RuntimeError(f"oracle batch_insert failed: {r}")


def sort_rows(P: Points, G: Graph, tie_mode=TIE_ID, nthreads=0):
    lib().orc_sort_rows(C.byref(P.s), C.byref(G.s), tie_mode, nthreads)


def ground_truth(P: Points, Q: np.ndarray, k, nthreads=0):
    Q = np.ascontiguousarray(Q)
    nq = Q.shape[0]
    ids = np.empty((nq, k), np.uint32); d = np.empty((nq, k), np.float64)
    assert lib().orc_ground_truth(C.byref(P.s), Q.ctypes.data, nq, k, ids.ctypes.data, d.ctypes.data,
                                  nthreads) == 0
    return ids, d


def quantized_mips(codes: np.ndarray, queries: np.ndarray, max_coord, min_coord) -> np.ndarray:
    """Quantized_Mips_Point<T>::distance of every (query, code) pair, nq x n;
    float32 queries are Mips_Point<float> queries, code-typed ones are
    quantized queries."""
    codes = np.ascontiguousarray(codes)
    queries = np.ascontiguousarray(queries)
    assert codes.shape[1] == queries.shape[1]
    qq = int(queries.dtype != np.float32)
    out = np.empty((queries.shape[0], codes.shape[0]), np.float32)
    r = lib().orc_quantized_mips(DTYPES[codes.dtype.name], codes.shape[1], max_coord, min_coord, queries.ctypes.data,
                                 qq, queries.shape[0], codes.ctypes.data, codes.shape[0], out.ctypes.data)
    assert r == 0
    return out
