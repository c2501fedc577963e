"""GPU beam search (libddg.so through the C-ABI) against the reference's own
search outputs (golden vectors) and against the oracle.  Bit-exact: ids,
distance bits, dist_cmps and |visited|."""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def _index(X, d, G, metric="l2"):
    from distributed_diskann_amd import Index
    idx = Index(X, d, metric, G.shape[1] - 1)
    idx.set_graph(G)
    return idx


@pytest.mark.parametrize("Ls", [10, 20, 50, 100, 200])
def test_search_reference_xs(xs, Ls):
    print(f"[gpu] search random-xs Ls={Ls}", flush=True)
    z = xs["z"]
    idx = _index(xs["X"], xs["d"], xs["G"])
    ids, dists, cmps, vis = idx.search(xs["Q"], 10, Ls, 1.35)
    assert np.array_equal(ids, z[f"ids_{Ls}"])
    assert np.array_equal(dists.view(np.uint32), z[f"dists_{Ls}"].view(np.uint32))
    assert np.array_equal(cmps, z[f"cmps_{Ls}"])
    assert np.array_equal(vis, z[f"visited_{Ls}"])


@pytest.mark.parametrize("Ls", [10, 50, 200])
def test_search_reference_n0(n0, Ls):
    print(f"[gpu] search n0 Ls={Ls}", flush=True)
    z = n0["z"]
    idx = _index(n0["X"], n0["d"], n0["G"])
    ids, dists, cmps, vis = idx.search(n0["Q"], 10, Ls, 1.35)
    assert np.array_equal(ids, z[f"ids_{Ls}"])
    assert np.array_equal(dists.view(np.uint32), z[f"dists_{Ls}"].view(np.uint32))
    assert np.array_equal(cmps, z[f"cmps_{Ls}"])
    assert np.array_equal(vis, z[f"visited_{Ls}"])


@pytest.mark.parametrize("Ls", [300, 1000])
def test_search_large_beams(oracle, xs, Ls):
    """Beams of the reference's sweep beyond 256: the visited filter (up to
    2^18 words) lives in global memory."""
    print(f"[gpu] search large beam Ls={Ls}", flush=True)
    idx = _index(xs["X"], xs["d"], xs["G"])
    Q = xs["Q"][:200]
    ids, dists, cmps, vis = idx.search(Q, 10, Ls, 1.35)
    P = oracle.Points(xs["X"], xs["d"])
    oi, od, oc, ov = oracle.search_all(P, oracle.Graph(xs["G"]), Q, 10, Ls, cut=1.35)
    assert np.array_equal(ids, oi)
    assert np.array_equal(dists.view(np.uint32), od.view(np.uint32))
    assert np.array_equal(cmps, oc) and np.array_equal(vis, ov)


@pytest.mark.parametrize("Ls", [50, 120])
def test_autotune_search(oracle, xs, Ls):
    """The autotuner only picks among result-preserving launch shapes: after
    tuning, searches still equal the oracle bit for bit."""
    print(f"[gpu] autotune Ls={Ls}", flush=True)
    idx = _index(xs["X"], xs["d"], xs["G"])
    info = idx.autotune_search(xs["Q"], 10, Ls)
    print(f"[gpu] autotune Ls={Ls}: {info}", flush=True)
    assert info["tried"] >= 4
    assert info["tuned_ms"] <= info["default_ms"]
    ids, dists, cmps, vis = idx.search(xs["Q"], 10, Ls, 1.35)
    P = oracle.Points(xs["X"], xs["d"])
    oi, od, oc, ov = oracle.search_all(P, oracle.Graph(xs["G"]), xs["Q"], 10, Ls, cut=1.35)
    assert np.array_equal(ids, oi)
    assert np.array_equal(dists.view(np.uint32), od.view(np.uint32))
    assert np.array_equal(cmps, oc) and np.array_equal(vis, ov)


def test_autotune_process_table(oracle, xs):
    """A tuned shape is kept for the whole process: a second handle over the
    same kind of index (dtype, metric, dimensions, degree bound) starts from
    it, with unchanged results; a handle with another degree bound does not."""
    from distributed_diskann_amd import Index
    print("[gpu] autotune process table", flush=True)
    Ls = 150  # a filter class no other test tunes
    a = _index(xs["X"], xs["d"], xs["G"])
    assert a.search_shape(Ls)["source"] == "default"
    info = a.autotune_search(xs["Q"], 10, Ls)
    sa = a.search_shape(Ls)
    assert sa["source"] == "handle"
    assert (sa["waves_per_block"], sa["blocks_per_cu"]) == (info["waves_per_block"], info["blocks_per_cu"])
    b = _index(xs["X"], xs["d"], xs["G"])
    sb = b.search_shape(Ls)
    print(f"[gpu] tuned {sa}, second handle {sb}", flush=True)
    assert sb["source"] == "process"
    assert {k: v for k, v in sb.items() if k != "source"} == {k: v for k, v in sa.items() if k != "source"}
    ids, dists, cmps, vis = b.search(xs["Q"], 10, Ls, 1.35)
    P = oracle.Points(xs["X"], xs["d"])
    oi, od, oc, ov = oracle.search_all(P, oracle.Graph(xs["G"]), xs["Q"], 10, Ls, cut=1.35)
    assert np.array_equal(ids, oi) and np.array_equal(dists.view(np.uint32), od.view(np.uint32))
    assert np.array_equal(cmps, oc) and np.array_equal(vis, ov)
    R2 = xs["G"].shape[1] - 1 + 8
    c = Index(xs["X"], xs["d"], "l2", R2)
    assert c.search_shape(Ls)["source"] == "default"
    for h in (a, b, c):
        h.close()


@pytest.mark.parametrize("stage", [2, 9])
@pytest.mark.parametrize("dtype,metric,d", [("float32", "l2", 200), ("float32", "mips", 150)])
def test_search_staged_rows(oracle, dtype, metric, d, stage):
    """Long rows gathered into LDS (stage rows per round trip): same bits as
    one lane per row, for the native and the reduced traversals."""
    print(f"[gpu] staged rows {metric} d={d} stage={stage}", flush=True)
    import ctypes as C
    from distributed_diskann_amd._lib import lib
    Xa, Qa, P, G = _small_case(oracle, dtype, metric, 3000, d, seed=d + stage)
    idx = _index(Xa, d, G.rows, metric)
    base = {m: idx.search(Qa, 10, 60, 1.35, trav=m) for m in ("native", "bf16", "fp8")}
    f = lib().ddg_internal_set_shape
    f.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32]
    f.restype = C.c_int
    for t, m in enumerate(("native", "bf16", "fp8")):
        assert f(idx.h, t, 60, 1, 4, 2, stage) == 0
        ids, dists, cmps, vis = idx.search(Qa, 10, 60, 1.35, trav=m)
        b = base[m]
        assert np.array_equal(ids, b[0]) and np.array_equal(dists.view(np.uint32), b[1].view(np.uint32))
        assert np.array_equal(cmps, b[2]) and np.array_equal(vis, b[3])
    oi, od, oc, ov = oracle.search_all(P, G, Qa, 10, 60, cut=1.35)
    assert np.array_equal(base["native"][0], oi) and np.array_equal(base["native"][2], oc)


@pytest.mark.parametrize("stage", [3, 8])
def test_search_staged_rows_low_degree(stage):
    """The 10M x 200 MIPS workload's distribution at 30,000 points (alpha
    1.0, R=64: about 5 neighbors per row, so most expansions have a few
    candidates): staged rows (one lane per row) give the same bits, counts and
    visits as rows split over lanes."""
    print(f"[gpu] staged rows, MIPS mixture, low degree, stage={stage}", flush=True)
    import ctypes as C
    from distributed_diskann_amd import Index, datasets
    from distributed_diskann_amd._lib import lib
    X, Q, w = datasets.generate("mips", n=30_000, nq=400)
    idx = Index(X, w.d, w.metric, w.R)
    idx.build(R=w.R, L=w.L, alpha=w.alpha)
    base = {m: idx.search(Q, 10, 100, 1.35, trav=m, rerank=50) for m in ("bf16", "fp8")}
    f = lib().ddg_internal_set_shape
    f.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32]
    f.restype = C.c_int
    for t, m in ((1, "bf16"), (2, "fp8")):
        assert f(idx.h, t, 100, 1, 4, 2, stage) == 0
        ids, dists, cmps, vis = idx.search(Q, 10, 100, 1.35, trav=m, rerank=50)
        b = base[m]
        assert np.array_equal(ids, b[0]) and np.array_equal(dists.view(np.uint32), b[1].view(np.uint32))
        assert np.array_equal(cmps, b[2]) and np.array_equal(vis, b[3])


def test_search_per_query_beam(oracle, xs):
    print("[gpu] per-query beam", flush=True)
    rng = np.random.default_rng(7)
    beams = np.array([10, 20, 50, 100, 200], np.uint32)[rng.integers(0, 5, xs["Q"].shape[0])]
    idx = _index(xs["X"], xs["d"], xs["G"])
    ids, dists, cmps, vis = idx.search(xs["Q"], 10, 200, 1.35, per_query_beam=beams)
    P = oracle.Points(xs["X"], xs["d"])
    oi, od, oc, ov = oracle.search_all(P, oracle.Graph(xs["G"]), xs["Q"], 10, 200, cut=1.35, per_query_beam=beams)
    assert np.array_equal(ids, oi)
    assert np.array_equal(dists.view(np.uint32), od.view(np.uint32))
    assert np.array_equal(cmps, oc) and np.array_equal(vis, ov)


def _small_case(oracle, dtype, metric, n, d, seed, R=24, L=40, alpha=1.2, nq=300):
    from distributed_diskann_amd.formats import align_rows
    rng = np.random.default_rng(seed)
    if dtype == "uint8":
        C = rng.uniform(0, 200, (20, d))
        X = np.clip(np.rint(C[rng.integers(0, 20, n)] + rng.normal(0, 15, (n, d))), 0, 255).astype(np.uint8)
        Q = np.clip(np.rint(C[rng.integers(0, 20, nq)] + rng.normal(0, 15, (nq, d))), 0, 255).astype(np.uint8)
    elif dtype == "int8":
        C = rng.normal(0, 30, (20, d))
        X = np.clip(np.rint(C[rng.integers(0, 20, n)] + rng.normal(0, 12, (n, d))), -128, 127).astype(np.int8)
        Q = np.clip(np.rint(C[rng.integers(0, 20, nq)] + rng.normal(0, 12, (nq, d))), -128, 127).astype(np.int8)
    else:
        X = rng.standard_normal((n, d)).astype(np.float32)
        Q = rng.standard_normal((nq, d)).astype(np.float32)
    Xa, Qa = align_rows(X), align_rows(Q)
    P = oracle.Points(Xa, d, metric)
    G = oracle.build(P, R, L, alpha, tie_mode=oracle.TIE_ID)
    return Xa, Qa, P, G


@pytest.mark.parametrize("dtype,metric,d", [("uint8", "l2", 128), ("int8", "l2", 100), ("float32", "mips", 64),
                                            ("uint8", "mips", 32), ("float32", "l2", 200), ("int8", "mips", 48)])
def test_search_vs_oracle_dtypes(oracle, dtype, metric, d):
    print(f"[gpu] search {dtype} {metric} d={d}", flush=True)
    Xa, Qa, P, G = _small_case(oracle, dtype, metric, 3000, d, seed=d)
    idx = _index(Xa, d, G.rows, metric)
    for Ls in (10, 40, 100):
        ids, dists, cmps, vis = idx.search(Qa, 10, Ls, 1.35)
        oi, od, oc, ov = oracle.search_all(P, G, Qa, 10, Ls, cut=1.35)
        assert np.array_equal(ids, oi)
        assert np.array_equal(dists.view(np.uint32), od.view(np.uint32))
        assert np.array_equal(cmps, oc) and np.array_equal(vis, ov)


@pytest.mark.parametrize("dtype,metric", [("float32", "mips"), ("int8", "mips"), ("float32", "l2"), ("uint8", "l2")])
def test_search_zero_distances(oracle, dtype, metric):
    """Many exact zero distances: inner products of vectors with disjoint
    supports (every zero is -0.0) and, for L2, base rows that repeat the
    queries (+0.0).  Candidates are ranked by integer keys of the distance
    bits, so zeros and their ties must come out exactly as the float order."""
    print(f"[gpu] search zero distances {dtype} {metric}", flush=True)
    from distributed_diskann_amd.formats import align_rows
    rng = np.random.default_rng(17)
    n, d, nq = 3000, 32, 200
    if dtype == "float32":
        X = rng.standard_normal((n, d)).astype(np.float32)
        Q = rng.standard_normal((nq, d)).astype(np.float32)
    else:
        lo, hi = (0, 4) if dtype == "uint8" else (-3, 4)
        X = rng.integers(lo, hi, (n, d)).astype(dtype)
        Q = rng.integers(lo, hi, (nq, d)).astype(dtype)
    if metric == "mips":
        X[: n // 2, : d // 2] = 0  # half the base is orthogonal to every query
        Q[:, d // 2:] = 0
    else:
        X[:nq] = Q  # exact copies of the queries
        X[nq:2 * nq] = Q
    Xa, Qa = align_rows(X), align_rows(Q)
    P = oracle.Points(Xa, d, metric)
    G = oracle.build(P, 16, 32, 1.2 if metric == "l2" else 1.0, tie_mode=oracle.TIE_ID)
    idx = _index(Xa, d, G.rows, metric)
    for Ls in (10, 50):
        ids, dists, cmps, vis = idx.search(Qa, 10, Ls, 1.35)
        oi, od, oc, ov = oracle.search_all(P, G, Qa, 10, Ls, cut=1.35)
        assert np.array_equal(ids, oi)
        assert np.array_equal(dists.view(np.uint32), od.view(np.uint32))
        assert np.array_equal(cmps, oc) and np.array_equal(vis, ov)


def test_search_limits_and_starts(oracle, xs):
    print("[gpu] limits / degree limits / starts", flush=True)
    idx = _index(xs["X"], xs["d"], xs["G"])
    P = oracle.Points(xs["X"], xs["d"])
    G = oracle.Graph(xs["G"])
    for limit, dl, starts in [(15, 32, (0,)), (10000, 9, (0,)), (10000, 32, (5, 9000, 77)), (40, 16, (3, 4))]:
        ids, dists, cmps, vis = idx.search(xs["Q"], 10, 50, 1.35, limit=limit, degree_limit=dl, starts=starts)
        oi, od, oc, ov = oracle.search_all(P, G, xs["Q"], 10, 50, cut=1.35, limit=limit, degree_limit=dl,
                                           starts=starts)
        assert np.array_equal(ids, oi) and np.array_equal(cmps, oc) and np.array_equal(vis, ov)
        assert np.array_equal(dists.view(np.uint32), od.view(np.uint32))


def test_search_duplicate_edges(oracle, xs):
    # append_neighbors allows duplicate ids in a row; the lossy filter and the
    # multiset merge must then behave exactly as the reference's
    print("[gpu] duplicate edges", flush=True)
    G = xs["G"].copy()
    rng = np.random.default_rng(1)
    for i in rng.integers(0, G.shape[0], 3000):
        deg = G[i, 0]
        if 2 <= deg < 32:
            G[i, 1 + deg] = G[i, 1 + rng.integers(0, deg)]
            G[i, 0] = deg + 1
    idx = _index(xs["X"], xs["d"], G)
    P = oracle.Points(xs["X"], xs["d"])
    for Ls in (10, 50):
        ids, dists, cmps, vis = idx.search(xs["Q"], 10, Ls, 1.35)
        oi, od, oc, ov = oracle.search_all(P, oracle.Graph(G), xs["Q"], 10, Ls, cut=1.35)
        assert np.array_equal(ids, oi) and np.array_equal(cmps, oc) and np.array_equal(vis, ov)


def test_search_wide_rows(oracle):
    # R = 96: more than one wave of neighbors and candidates per expansion,
    # with duplicate edges, against the oracle at a small and a large beam
    from distributed_diskann_amd import Index
    from distributed_diskann_amd.formats import align_rows
    print("[gpu] R=96 rows", flush=True)
    rng = np.random.default_rng(5)
    n, d, R = 1500, 16, 96
    X = align_rows(rng.standard_normal((n, d)).astype(np.float32))
    Q = align_rows(rng.standard_normal((64, d)).astype(np.float32))
    P = oracle.Points(X, d)
    G = oracle.build(P, R, 128, 1.2, tie_mode=oracle.TIE_ID).rows.copy()
    assert G[:, 0].max() > 64
    for i in rng.integers(0, n, 400):
        deg = G[i, 0]
        if 2 <= deg < R:
            G[i, 1 + deg] = G[i, 1 + rng.integers(0, deg)]
            G[i, 0] = deg + 1
    idx = Index(X, d, "l2", R)
    idx.set_graph(G)
    for Ls in (40, 150):
        ids, dists, cmps, vis = idx.search(Q, 10, Ls, 1.35)
        oi, od, oc, ov = oracle.search_all(P, oracle.Graph(G), Q, 10, Ls, cut=1.35)
        assert np.array_equal(ids, oi) and np.array_equal(cmps, oc) and np.array_equal(vis, ov)
        assert np.array_equal(dists.view(np.uint32), od.view(np.uint32))


def test_insert_traces_reference(xs):
    print("[gpu] build-mode traces", flush=True)
    z = xs["z"]
    idx = _index(xs["X"], xs["d"], xs["G"])
    lists, cmps = idx.insert_search(z["ins_ids"], 50, 0)
    assert np.array_equal(cmps, z["ins_cmps"])
    off = 0
    for i, (vi, vd) in enumerate(lists):
        vl = int(z["ins_vlen"][i])
        assert len(vi) == vl
        order = np.lexsort((vi, vd))  # the reference keeps visited sorted by (distance, id)
        assert np.array_equal(vi[order], z["ins_vis_ids"][off:off + vl])
        assert np.array_equal(vd[order].view(np.uint32), z["ins_vis_dists"][off:off + vl].view(np.uint32))
        off += vl


def test_search_errors(xs):
    from distributed_diskann_amd import DdgError
    idx = _index(xs["X"], xs["d"], xs["G"])
    with pytest.raises(DdgError):
        idx.search(xs["Q"][:4], 20, 10, 1.35)  # k > beam (beamSearch.h:327-331)
    with pytest.raises(DdgError):
        idx.search(xs["Q"][:4], 10, 20, 1.35, starts=(10**6,))
    with pytest.raises(DdgError):
        idx.search(xs["Q"][:4], 10, 20, 0.9)


def test_ground_truth_vs_oracle(oracle, xs):
    print("[gpu] ground truth", flush=True)
    idx = _index(xs["X"], xs["d"], xs["G"])
    ids, d = idx.ground_truth(xs["Q"][:300], 10)
    oi, od = oracle.ground_truth(oracle.Points(xs["X"], xs["d"]), xs["Q"][:300], 10)
    from test_gpu_dense import gt_match
    assert gt_match(ids, d, oi, od, xs["X"], xs["Q"][:300], xs["d"], "l2")
    for dtype, metric, dd in [("uint8", "l2", 128), ("int8", "mips", 100)]:
        Xa, Qa, P, G = _small_case(oracle, dtype, metric, 2000, dd, seed=3)
        idx2 = _index(Xa, dd, G.rows, metric)
        ids, d = idx2.ground_truth(Qa, 20)
        oi, od = oracle.ground_truth(P, Qa, 20)
        assert np.array_equal(ids, oi) and np.array_equal(d, od)


@pytest.mark.parametrize("order", ["pre", "start", "0"])
def test_search_query_order(oracle, monkeypatch, order):
    """The order a launch takes its queries in (by default the largest beam
    first, and for a long uniform beam, here from 100 on, the descending best
    distance of a short pre-pass search; by distance to the start point with
    DDG_ORDER=start; none with DDG_ORDER=0) changes no result: 12,000 queries
    (more than the resident waves) bit-exact against the oracle, with uniform
    and per-query beams."""
    from distributed_diskann_amd.formats import align_rows
    print(f"[gpu] query order {order}", flush=True)
    if order == "pre":
        monkeypatch.delenv("DDG_ORDER", raising=False)
        monkeypatch.setenv("DDG_ORDER_MIN_BEAM", "100")
    else:
        monkeypatch.setenv("DDG_ORDER", order)
    rng = np.random.default_rng(77)
    d = 20
    X = align_rows(rng.standard_normal((20000, d)).astype(np.float32))
    Q = align_rows(rng.standard_normal((12000, d)).astype(np.float32))
    P = oracle.Points(X, d)
    G = oracle.build(P, 16, 32, 1.2, tie_mode=oracle.TIE_ID)
    idx = _index(X, d, G.rows)
    beams = rng.choice(np.array([10, 20, 50, 100], np.uint32), Q.shape[0])
    for pqb in (None, beams):
        ids, dists, cmps, vis = idx.search(Q, 10, 100, 1.35, per_query_beam=pqb)
        oi, od, oc, ov = oracle.search_all(P, G, Q, 10, 100, cut=1.35, per_query_beam=pqb)
        assert np.array_equal(ids, oi)
        assert np.array_equal(dists.view(np.uint32), od.view(np.uint32))
        assert np.array_equal(cmps, oc) and np.array_equal(vis, ov)
    idx.close()


@pytest.mark.parametrize("dtype,metric,trav", [("float32", "mips", "fp8"), ("float32", "l2", "bf16"),
                                               ("uint8", "l2", "native"), ("float32", "mips", "native")])
def test_search_prepass_order_identical(oracle, monkeypatch, dtype, metric, trav):
    """Uniform-beam launches with more queries than resident waves take their
    queries in the order of a short pre-pass search (beam 10, k = 1, 8 visits,
    re-ranking its best entry in the reduced modes) when the beam is 140 or
    more; every output equals the launch in the queries' own order
    (DDG_ORDER=0), at beams above and below that threshold."""
    print(f"[gpu] pre-pass order {dtype} {metric} {trav}", flush=True)
    torch = pytest.importorskip("torch")
    d = 48
    X, Q, P, G = _small_case(oracle, dtype, metric, 10000, d, seed=11, nq=12000)
    idx = _index(X, d, G.rows, metric)
    if trav != "native":
        idx.prepare_traversal(trav)
    dev = torch.device("cuda:0")
    Qd = torch.from_numpy(Q).to(dev)
    nq = Q.shape[0]
    starts = torch.zeros(1, dtype=torch.int32, device=dev)
    rr = 100 if trav != "native" else 0

    def run(Ls):
        ids = torch.empty((nq, 10), dtype=torch.int32, device=dev)
        dd = torch.empty((nq, 10), dtype=torch.float32, device=dev)
        cm = torch.empty(nq, dtype=torch.int64, device=dev)
        vs = torch.empty(nq, dtype=torch.int32, device=dev)
        idx.search_device(Qd, 10, Ls, ids, dd, cm, vs, cut=1.35, starts=starts, trav=trav, rerank=rr)
        torch.cuda.synchronize()
        return [t.cpu().numpy() for t in (ids, dd.view(torch.int32), cm, vs)]

    for Ls in (50, 150):
        monkeypatch.delenv("DDG_ORDER", raising=False)
        got = run(Ls)
        monkeypatch.setenv("DDG_ORDER", "0")
        want = run(Ls)
        for g, w in zip(got, want):
            assert np.array_equal(g, w)
    if trav == "native" and dtype == "uint8":
        # three start points and caller-given query ids (the self-skip id),
        # which the pre-pass takes too; the host-buffer entry point as well
        starts = torch.tensor([0, 17, 4242], dtype=torch.int32, device=dev)
        qids = np.random.default_rng(5).integers(0, 10000, nq).astype(np.uint32)
        Qid = torch.from_numpy(qids.view(np.int32)).to(dev)

        def run_ids(Ls):
            ids = torch.empty((nq, 10), dtype=torch.int32, device=dev)
            dd = torch.empty((nq, 10), dtype=torch.float32, device=dev)
            cm = torch.empty(nq, dtype=torch.int64, device=dev)
            vs = torch.empty(nq, dtype=torch.int32, device=dev)
            idx.search_device(Qd, 10, Ls, ids, dd, cm, vs, cut=1.35, starts=starts, query_ids=Qid)
            torch.cuda.synchronize()
            return [t.cpu().numpy() for t in (ids, dd.view(torch.int32), cm, vs)]

        monkeypatch.delenv("DDG_ORDER", raising=False)
        got = run_ids(150)
        host = idx.search(Q, 10, 150, 1.35, starts=np.array([0, 17, 4242], np.uint32), query_ids=qids)
        monkeypatch.setenv("DDG_ORDER", "0")
        want = run_ids(150)
        for g, w in zip(got, want):
            assert np.array_equal(g, w)
        assert np.array_equal(host[0].view(np.uint32), want[0].view(np.uint32))
        assert np.array_equal(host[1].view(np.uint32), want[1].view(np.uint32))
        assert np.array_equal(host[2], want[2]) and np.array_equal(host[3], want[3])
    idx.close()
