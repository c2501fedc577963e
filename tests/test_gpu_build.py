"""GPU index build (prefix-doubling batch insert, robustPrune, reverse edges,
final row sort) against the oracle.  The GPU orders prune candidates by
(distance, id); the oracle's TIE_ID mode does the same, so the graphs must be
identical row for row.  Against the reference's own tie order (std::sort by
distance only) the contract is recall within 0.005."""
import numpy as np
import pytest

from conftest import canon
from test_gpu_search import _small_case

pytestmark = pytest.mark.gpu


def _gpu_build(Xa, d, metric, R, **kw):
    from distributed_diskann_amd import Index
    idx = Index(Xa, d, metric, R)
    secs = idx.build(R=R, **kw)
    return idx, idx.get_graph(), secs


@pytest.mark.parametrize("dtype,metric,d,alpha", [("float32", "l2", 20, 1.2), ("uint8", "l2", 128, 1.2),
                                                  ("int8", "l2", 100, 1.2), ("float32", "mips", 64, 1.0),
                                                  ("uint8", "mips", 32, 1.0)])
def test_build_identical_to_oracle(oracle, dtype, metric, d, alpha):
    print(f"[gpu] build {dtype} {metric} d={d}", flush=True)
    Xa, Qa, P, _ = _small_case(oracle, dtype, metric, 4000, d, seed=11, R=24, L=40)
    G = oracle.build(P, 24, 40, alpha, tie_mode=oracle.TIE_ID)
    idx, rows, secs = _gpu_build(Xa, d, metric, 24, L=40, alpha=alpha)
    assert np.array_equal(canon(rows), canon(G.rows))


@pytest.mark.parametrize("dtype,metric,d,hi,alpha", [("uint8", "l2", 128, 3, 1.2), ("uint8", "l2", 64, 2, 1.0),
                                                     ("int8", "mips", 32, 2, 1.0), ("int8", "l2", 256, 128, 1.5),
                                                     ("uint8", "l2", 160, 4, 1.2), ("uint8", "mips", 200, 3, 1.0),
                                                     ("int8", "l2", 100, 3, 0.95), ("uint8", "l2", 256, 256, 1.2),
                                                     ("uint8", "l2", 64, 3, 0.0)])
def test_build_integer_ties(oracle, dtype, metric, d, hi, alpha):
    """Low-entropy integer rows: many equal distances, and prune tests that
    land exactly on alpha * d(a, b) == d(p, b) (the matrix-core prune decides
    them by integer thresholds).  Row widths 64, 128, 192 and 256 bytes (one
    to four K steps per tile); the last case reaches the largest uint8
    distances (255^2 x 256, just below 2^24); alpha = 0 takes the pairwise
    prune instead of the matrix-core one (every test is 0 <= d(p, b))."""
    print(f"[gpu] build integer ties {dtype} {metric} d={d} values<{hi} alpha={alpha}", flush=True)
    from distributed_diskann_amd.formats import align_rows
    rng = np.random.default_rng(d + hi)
    lo = 0 if dtype == "uint8" else -hi
    Xa = align_rows(rng.integers(lo, hi, (3000, d)).astype(dtype))
    G = oracle.build(oracle.Points(Xa, d, metric), 16, 32, alpha, tie_mode=oracle.TIE_ID)
    idx, rows, secs = _gpu_build(Xa, d, metric, 16, L=32, alpha=alpha)
    assert np.array_equal(canon(rows), canon(G.rows))


def test_build_hub_identical(oracle):
    """A star: points on a shell around a centre point, which is among every
    point's nearest neighbours, so its reverse-edge lists in one batch run to
    thousands of candidates and take the queued one-workgroup-per-task pass,
    writing the graph from bump-allocated scratch."""
    print("[gpu] build star (hub reverse lists) 20000 x 16", flush=True)
    from distributed_diskann_amd.formats import align_rows
    rng = np.random.default_rng(21)
    n, d = 20000, 16
    dirs = rng.standard_normal((n, d))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    X = (dirs * (1.0 + 0.05 * rng.standard_normal((n, 1)))).astype(np.float32)
    X[0] = 0.0
    Xa = align_rows(X)
    G = oracle.build(oracle.Points(Xa, d), 16, 32, 1.2, tie_mode=oracle.TIE_ID)
    idx, rows, secs = _gpu_build(Xa, d, "l2", 16, L=32, alpha=1.2)
    assert np.array_equal(canon(rows), canon(G.rows))


def test_build_hub_windows(oracle):
    """Tight clusters at staggered distances around a centre point (cluster k
    at radius 1 + 0.02 k).  Every point links to the centre, so its reverse
    lists hold thousands of candidates ordered cluster by cluster, and each
    selection prunes only its own cluster: the hub sweep needs more than one
    window of the nearest candidates, and a later window's candidates from a
    cluster that straddles the window bound are pruned by a point selected in
    an earlier window."""
    print("[gpu] build clustered star (windowed hub sweep) 18001 x 16", flush=True)
    from distributed_diskann_amd.formats import align_rows
    rng = np.random.default_rng(33)
    K, per, d = 60, 300, 16
    dirs = rng.standard_normal((K, d))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    cent = dirs * (1.0 + 0.02 * np.arange(K))[:, None]
    X = np.concatenate([np.zeros((1, d)), np.repeat(cent, per, 0) + 0.01 * rng.standard_normal((K * per, d))])
    X = X[np.concatenate([[0], 1 + rng.permutation(K * per)])].astype(np.float32)
    Xa = align_rows(X)
    G = oracle.build(oracle.Points(Xa, d), 16, 32, 1.2, tie_mode=oracle.TIE_ID)
    idx, rows, secs = _gpu_build(Xa, d, "l2", 16, L=32, alpha=1.2)
    assert np.array_equal(canon(rows), canon(G.rows))


def test_build_hub_ties(oracle):
    """A centre point (the zero row, id 0) and 0/1 rows with exactly four
    ones in 64 columns: the centre is at distance 4 from every point and
    prunes its other candidates, so every point links to it, and its reverse
    lists hold thousands of candidates at one distance.  More than a window's
    worth of equal keys: the hub sweep runs over the list in place, in id
    order."""
    print("[gpu] build tied star (hub list of equal distances) 12001 x 64 uint8", flush=True)
    from distributed_diskann_amd.formats import align_rows
    rng = np.random.default_rng(34)
    n, d = 12000, 64
    X = np.zeros((n + 1, d), np.uint8)
    for i in range(1, n + 1):
        X[i, rng.choice(d, 4, replace=False)] = 1
    Xa = align_rows(X)
    G = oracle.build(oracle.Points(Xa, d), 16, 32, 1.2, tie_mode=oracle.TIE_ID)
    idx, rows, secs = _gpu_build(Xa, d, "l2", 16, L=32, alpha=1.2)
    assert np.array_equal(canon(rows), canon(G.rows))


def test_build_n0_identical_and_recall(oracle, n0):
    print("[gpu] build n0 (10k x 20, R32 L50)", flush=True)
    from distributed_diskann_amd.index import recall_with_ties
    w = n0["w"]
    P = oracle.Points(n0["X"], w.d)
    G = oracle.build(P, w.R, w.L, w.alpha, tie_mode=oracle.TIE_ID)
    idx, rows, secs = _gpu_build(n0["X"], w.d, "l2", w.R, L=w.L, alpha=w.alpha)
    assert np.array_equal(canon(rows), canon(G.rows))
    gt_ids, gt_d = idx.ground_truth(n0["Q"], 10)
    ids, _, _, _ = idx.search(n0["Q"], 10, 50, 1.35)
    r_gpu = recall_with_ties(ids, gt_ids, gt_d, 10)
    # the reference's own tie order (graph fixture built like index.h)
    oi, _, _, _ = oracle.search_all(P, oracle.Graph(n0["G"]), n0["Q"], 10, 50, cut=1.35)
    r_ref = recall_with_ties(oi, gt_ids, gt_d, 10)
    assert r_gpu >= r_ref - 0.005


def test_build_random_xs_recall(oracle, xs):
    print("[gpu] build random-xs", flush=True)
    from distributed_diskann_amd.index import recall_with_ties
    idx, rows, secs = _gpu_build(xs["X"], xs["d"], "l2", 32, L=50, alpha=1.2)
    ids, _, _, _ = idx.search(xs["Q"], 10, 50, 1.35)
    r_gpu = recall_with_ties(ids, xs["gt_ids"], xs["gt_d"], 10)
    oi, _, _, _ = oracle.search_all(oracle.Points(xs["X"], xs["d"]), oracle.Graph(xs["G"]), xs["Q"], 10, 50,
                                    cut=1.35)
    r_ref = recall_with_ties(oi, xs["gt_ids"], xs["gt_d"], 10)
    assert r_gpu >= r_ref - 0.005


@pytest.mark.parametrize("kw", [dict(bsp=True, max_fraction=0.02, start_point=17, final_sort=False),
                                dict(bsp=True, max_fraction=1.0, max_build=2500, final_sort=False),
                                dict(two_pass=True), dict(final_sort=False, base=3.0)])
def test_build_variants(oracle, kw):
    print(f"[gpu] build variant {kw}", flush=True)
    Xa, Qa, P, _ = _small_case(oracle, "float32", "l2", 4000, 16, seed=5)
    G = oracle.build(P, 24, 40, 1.2, tie_mode=oracle.TIE_ID, **kw)
    idx, rows, secs = _gpu_build(Xa, 16, "l2", 24, L=40, alpha=1.2, **kw)
    assert np.array_equal(canon(rows), canon(G.rows))


def test_batch_insert_subset(oracle):
    print("[gpu] batch_insert subset", flush=True)
    from distributed_diskann_amd import Index
    Xa, Qa, P, _ = _small_case(oracle, "uint8", "l2", 3000, 64, seed=9)
    ins = np.random.default_rng(2).permutation(3000)[:1800].astype(np.uint32)
    G = oracle.Graph(np.zeros((3000, 25), np.uint32))
    oracle.batch_insert(P, G, ins, 1.2, 40, base=2.0, max_fraction=0.02, start=int(ins[0]),
                        tie_mode=oracle.TIE_ID)
    idx = Index(Xa, 64, "l2", 24)
    idx.batch_insert(ins, 1.2, 40, start_point=int(ins[0]), base=2.0, max_fraction=0.02)
    assert np.array_equal(canon(idx.get_graph()), canon(G.rows))


@pytest.mark.parametrize("m", [1800, 9000])
def test_batch_insert_random_order(oracle, m):
    # random_order = true inserts in parlay::random_permutation's order
    # (index.h:237-241), below and above its sequential threshold (8192)
    print(f"[gpu] batch_insert random_order m={m}", flush=True)
    from distributed_diskann_amd import Index
    n = 10000
    Xa, Qa, P, _ = _small_case(oracle, "uint8", "l2", n, 32, seed=11)
    ins = np.random.default_rng(3).permutation(n)[:m].astype(np.uint32)
    perm = oracle.random_permutation(m)
    G = oracle.Graph(np.zeros((n, 25), np.uint32))
    oracle.batch_insert(P, G, ins[perm], 1.2, 40, base=2.0, max_fraction=0.02, start=int(ins[0]),
                        tie_mode=oracle.TIE_ID)
    idx = Index(Xa, 32, "l2", 24)
    idx.batch_insert(ins, 1.2, 40, start_point=int(ins[0]), base=2.0, max_fraction=0.02, random_order=True)
    assert np.array_equal(canon(idx.get_graph()), canon(G.rows))


def test_robust_prune_vs_oracle(oracle):
    print("[gpu] robust_prune", flush=True)
    from distributed_diskann_amd import Index
    Xa, Qa, P, G = _small_case(oracle, "int8", "l2", 3000, 100, seed=4)
    idx = Index(Xa, 100, "l2", 24)
    idx.set_graph(G.rows)
    rng = np.random.default_rng(0)
    pts = rng.integers(0, 3000, 200).astype(np.uint32)
    cands = [rng.integers(0, 3000, rng.integers(0, 600)).astype(np.uint32) for _ in pts]
    for add in (True, False):
        for alpha in (1.0, 1.2):
            got = idx.robust_prune(pts, cands, None, alpha, add, 24)
            for p, c, g in zip(pts, cands, got):
                o = oracle.robust_prune(P, G, int(p), c, None, alpha, add, 24, tie_mode=oracle.TIE_ID)
                assert list(g) == list(o)
    # with caller-supplied distances
    dd = [oracle.distances("int8", "l2", Xa[c], np.repeat(Xa[[p]], len(c), 0), 100) for p, c in zip(pts, cands)]
    got = idx.robust_prune(pts, cands, dd, 1.2, True, 24)
    for p, c, d_, g in zip(pts, cands, dd, got):
        o = oracle.robust_prune(P, G, int(p), c, d_, 1.2, True, 24, tie_mode=oracle.TIE_ID)
        assert list(g) == list(o)


def test_build_errors(oracle):
    from distributed_diskann_amd import DdgError, Index
    Xa, Qa, P, G = _small_case(oracle, "float32", "l2", 500, 16, seed=1)
    idx = Index(Xa, 16, "l2", 24)
    with pytest.raises(DdgError):
        idx.batch_insert([0, 1, 10**6], 1.2, 40)  # index.h:217-223
    with pytest.raises(DdgError):
        idx.build(R=48, L=40)  # R above the graph's max degree
    bad = np.zeros((500, 25), np.uint32)
    bad[3, 0] = 30
    with pytest.raises(DdgError):
        idx.set_graph(bad)


@pytest.mark.parametrize("n_pairs,n_keys", [(1, 10), (5000, 300), (70000, 4000), (300000, 1 << 20)])
def test_reverse_pair_grouping(n_pairs, n_keys):
    # the stable device radix sort + group boundaries of the reverse-edge step,
    # against numpy's stable argsort (parlay::group_by_key keeps input order)
    print(f"[gpu] pair grouping {n_pairs} pairs over {n_keys} keys", flush=True)
    import ctypes as C
    from distributed_diskann_amd import Index
    from distributed_diskann_amd._lib import lib
    rng = np.random.default_rng(n_pairs)
    idx = Index(np.zeros((n_keys, 16), np.float32), 16, "l2", 8)
    keys = rng.integers(0, n_keys, n_pairs).astype(np.uint32)
    vals = np.arange(n_pairs, dtype=np.uint32)
    ok = np.empty(n_pairs, np.uint32); ov = np.empty(n_pairs, np.uint32)
    ng = np.zeros(1, np.uint64); gk = np.empty(n_pairs, np.uint32); go = np.empty(n_pairs + 1, np.uint64)
    mg = np.zeros(1, np.uint32)
    f = lib().ddg_internal_group_pairs
    f.restype = C.c_int
    p = lambda a: C.c_void_p(a.ctypes.data)
    assert f(idx.h, p(keys), p(vals), C.c_uint64(n_pairs), p(ok), p(ov), p(ng), p(gk), p(go), p(mg)) == 0
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(ok, keys[order]) and np.array_equal(ov, vals[order])
    uk, starts, counts = np.unique(keys[order], return_index=True, return_counts=True)
    g = int(ng[0])
    assert g == len(uk) and np.array_equal(gk[:g], uk) and np.array_equal(go[:g], starts)
    assert go[g] == n_pairs and mg[0] == counts.max()


@pytest.mark.parametrize("vcap,graphs", [("8", "1"), ("24", "1"), ("8", "0")])
def test_build_visited_overflow_rerun(oracle, monkeypatch, vcap, graphs):
    """Visited-list slots far shorter than the insert searches' lists: every
    early round overflows inside its recorded graph (whose later kernels then
    change nothing), is rerun with slots twice as long, and the build still
    equals the oracle's row for row; the same launch by launch."""
    print(f"[gpu] build with visited slots starting at {vcap} (graphs {graphs})", flush=True)
    monkeypatch.setenv("DDG_BUILD_VCAP", vcap)
    monkeypatch.setenv("DDG_GRAPH", graphs)
    Xa, Qa, P, _ = _small_case(oracle, "float32", "l2", 6000, 24, seed=5, R=24, L=48)
    G = oracle.build(P, 24, 48, 1.2, tie_mode=oracle.TIE_ID)
    idx, rows, secs = _gpu_build(Xa, 24, "l2", 24, L=48, alpha=1.2)
    assert np.array_equal(canon(rows), canon(G.rows))
