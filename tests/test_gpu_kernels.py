"""Component parity of the build kernels, one launch mode at a time, against
the oracle: forward robustPrune over fixed-stride visited lists, the reverse
append-or-prune step writing the graph, candidate lists beyond the LDS
capacity (global scratch path), and the final per-row sort."""
import ctypes as C

import numpy as np
import pytest

from conftest import canon
from test_gpu_search import _small_case

pytestmark = pytest.mark.gpu


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _prune_hook(idx, pts, a_ids, a_dists, a_len, a_stride, a_off, alpha, Rsel, add, write_graph, append, max_cand):
    from distributed_diskann_amd._lib import lib
    f = lib().ddg_internal_prune
    f.restype = C.c_int
    npts = len(pts)
    out = np.zeros((npts, Rsel), np.uint32)
    ln = np.zeros(npts, np.uint32)
    st = f(idx.h, C.c_uint64(npts), _p(pts), _p(a_ids), _p(a_dists), _p(a_len), C.c_uint32(a_stride), _p(a_off),
           C.c_uint64(len(a_ids)), C.c_double(alpha), C.c_uint32(Rsel), C.c_int(add), C.c_int(write_graph),
           C.c_int(append), C.c_uint32(max_cand), _p(out), _p(ln))
    assert st == 0, st
    return out, ln


def _index(Xa, d, G, metric="l2"):
    from distributed_diskann_amd import Index
    idx = Index(Xa, d, metric, G.shape[1] - 1)
    idx.set_graph(G)
    return idx


@pytest.mark.parametrize("dtype,metric,d", [("float32", "l2", 20), ("uint8", "l2", 128), ("float32", "mips", 32)])
def test_forward_prune_fixed_stride(oracle, dtype, metric, d):
    print(f"[gpu] forward prune {dtype} {metric}", flush=True)
    Xa, Qa, P, G = _small_case(oracle, dtype, metric, 3000, d, seed=7)
    idx = _index(Xa, d, G.rows, metric)
    rng = np.random.default_rng(1)
    ntask, stride = 300, 144
    pts = rng.choice(3000, ntask, replace=False).astype(np.uint32)
    ids = np.zeros((ntask, stride), np.uint32)
    ds = np.zeros((ntask, stride), np.float32)
    lens = rng.integers(1, stride + 1, ntask).astype(np.uint32)
    for t in range(ntask):
        c = rng.choice(3000, lens[t], replace=False).astype(np.uint32)
        ids[t, :lens[t]] = c
        ds[t, :lens[t]] = oracle.distances(dtype, metric, Xa[c], np.repeat(Xa[[pts[t]]], len(c), 0), d)
    out, ln = _prune_hook(idx, pts, ids.ravel(), ds.ravel(), lens, stride, None, 1.2, 24, 1, 0, 0, stride + 24)
    for t in range(ntask):
        o = oracle.robust_prune(P, G, int(pts[t]), ids[t, :lens[t]], ds[t, :lens[t]], 1.2, True, 24,
                                tie_mode=oracle.TIE_ID)
        assert list(out[t, :ln[t]]) == list(o)


def test_reverse_append_or_prune(oracle):
    print("[gpu] reverse append / prune into the graph", flush=True)
    Xa, Qa, P, G = _small_case(oracle, "int8", "l2", 3000, 100, seed=8)
    idx = _index(Xa, 100, G.rows)
    rng = np.random.default_rng(2)
    targets = rng.choice(3000, 400, replace=False).astype(np.uint32)
    sizes = rng.integers(1, 60, 400)
    off = np.zeros(401, np.uint64)
    off[1:] = np.cumsum(sizes)
    S = [rng.integers(0, 3000, s).astype(np.uint32) for s in sizes]
    flat = np.concatenate(S)
    _prune_hook(idx, targets, flat, None, None, 0, off, 1.2, 24, 1, 1, 1, int(sizes.max()) + 24)
    got = canon(idx.get_graph())
    want = G.rows.copy()
    for t, s in zip(targets, S):
        row = want[t]
        if len(s) + row[0] <= 24:  # append_neighbors (graph.h:84-93)
            row[1 + row[0]:1 + row[0] + len(s)] = s
            row[0] += len(s)
        else:
            o = oracle.robust_prune(P, G, int(t), s, None, 1.2, True, 24, tie_mode=oracle.TIE_ID)
            row[0] = len(o)
            row[1:1 + len(o)] = o
    assert np.array_equal(got, canon(want))


@pytest.mark.parametrize("dtype,metric,d", [("float32", "l2", 20), ("uint8", "l2", 128), ("int8", "mips", 64)])
def test_prune_lists_beyond_lds(oracle, dtype, metric, d):
    """Lists of thousands of candidates (with repeated ids) take the queued
    pass, one workgroup per task."""
    print(f"[gpu] prune over candidate lists beyond the LDS capacity {dtype} {metric}", flush=True)
    Xa, Qa, P, G = _small_case(oracle, dtype, metric, 12000, d, seed=9)
    idx = _index(Xa, d, G.rows, metric)
    rng = np.random.default_rng(3)
    pts = rng.choice(12000, 6, replace=False).astype(np.uint32)
    sizes = np.array([5000, 7000, 100, 9000, 4100, 30])
    off = np.zeros(7, np.uint64)
    off[1:] = np.cumsum(sizes)
    S = [rng.integers(0, 12000, s).astype(np.uint32) for s in sizes]
    alpha = 1.2 if metric == "l2" else 1.0
    out, ln = _prune_hook(idx, pts, np.concatenate(S), None, None, 0, off, alpha, 24, 1, 0, 0, int(sizes.max()) + 24)
    for t in range(6):
        o = oracle.robust_prune(P, G, int(pts[t]), S[t], None, alpha, True, 24, tie_mode=oracle.TIE_ID)
        assert list(out[t, :ln[t]]) == list(o)


@pytest.mark.parametrize("dtype,metric,d", [("float32", "l2", 20), ("uint8", "l2", 128), ("int8", "mips", 64)])
def test_sort_rows(oracle, dtype, metric, d):
    print(f"[gpu] row sort {dtype} {metric}", flush=True)
    Xa, Qa, P, G = _small_case(oracle, dtype, metric, 3000, d, seed=12)
    rows = G.rows.copy()
    rng = np.random.default_rng(4)
    for i in range(3000):  # shuffle each list, with a duplicate id now and then
        k = rows[i, 0]
        rows[i, 1:1 + k] = rng.permutation(rows[i, 1:1 + k])
        if 1 <= k < 24 and i % 7 == 0:
            rows[i, 1 + k] = rows[i, 1]
            rows[i, 0] = k + 1
    idx = _index(Xa, d, rows, metric)
    idx.sort_rows()
    want = oracle.Graph(rows.copy())
    oracle.sort_rows(P, want, tie_mode=oracle.TIE_ID)
    assert np.array_equal(canon(idx.get_graph()), canon(want.rows))
