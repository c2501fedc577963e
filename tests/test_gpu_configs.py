"""Parity at the benchmark workloads' own sizes and parameters (GPU).

* 100,000 x 20 float32 L2, R=32 L=64 alpha=1.2: the whole GPU build equals
  the oracle's (distance, id)-ordered build row for row; the per-query-beam
  search of all 10,000 queries (beams uniform over {10,20,50,100,200}) is
  bit-exact against the oracle's searchAll with the same beams.
* 1,000,000 x 100 int8 L2, R=64 L=128 alpha=1.2: the whole GPU build equals
  the oracle's row for row (candidate lists beyond 256 entries, 24-bit radix
  keys, hipGraph rounds of up to 1M points); search of 1,000 queries
  bit-exact at Ls 40 and 128.
* 10,000,000 x 128 uint8 L2 and 10,000,000 x 200 float32 MIPS (alpha 1.0),
  R=64 L=128: the first 200,000 points built on the GPU equal the oracle's
  build of them; the full 10M build keeps every row within R, with ids in
  range and no self-loops, and a second build is byte-identical; the
  oracle's search of 200 queries over the GPU-built 10M graph equals the GPU
  search bit for bit (ids, distance bits, dist_cmps, |visited|); on the
  float32 set the BF16 and FP8 traversals return re-ranked distances equal
  to the native FP32 distance bits of the ids they return.

Oracle builds use (distance, id) order (TIE_ID), the GPU's tie order."""
import numpy as np
import pytest

from conftest import canon

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu] {msg}", flush=True)


def _build(X, d, metric, R, L, alpha):
    from distributed_diskann_amd import Index
    idx = Index(X, d, metric, R)
    secs = idx.build(R=R, L=L, alpha=alpha)
    return idx, idx.get_graph(), secs


def _invariants(rows, R):
    n = rows.shape[0]
    deg = rows[:, 0]
    assert deg.max() <= R and deg.min() >= 1
    mask = np.arange(R)[None, :] < deg[:, None]
    nb = rows[:, 1:]
    assert (nb[mask] < n).all()
    own = np.broadcast_to(np.arange(n, dtype=np.uint32)[:, None], nb.shape)
    assert not (mask & (nb == own)).any()


def _search_exact(oracle, idx, X, Q, d, metric, rows, Ls):
    ids, dists, cmps, vis = idx.search(Q, 10, Ls, 1.35)
    oi, od, oc, ov = oracle.search_all(oracle.Points(X, d, metric), oracle.Graph(rows), Q, 10, Ls, cut=1.35)
    assert np.array_equal(ids, oi)
    assert np.array_equal(dists.view(np.uint32), od.view(np.uint32))
    assert np.array_equal(cmps, oc) and np.array_equal(vis, ov)


# ---------------------------------------------------------------- 100K x 20
@pytest.fixture(scope="module")
def cfg_s(oracle):
    from distributed_diskann_amd import datasets
    X, Q, w = datasets.generate("s")
    _say("100K x 20: GPU build")
    idx, rows, _ = _build(X, w.d, w.metric, w.R, w.L, w.alpha)
    return dict(X=X, Q=Q, w=w, idx=idx, rows=rows)


@pytest.mark.timeout(600)
def test_s_full_build_identical(oracle, cfg_s):
    w = cfg_s["w"]
    _say("100K x 20: oracle build")
    G = oracle.build(oracle.Points(cfg_s["X"], w.d, w.metric), w.R, w.L, w.alpha, tie_mode=oracle.TIE_ID)
    assert np.array_equal(canon(cfg_s["rows"]), canon(G.rows))


@pytest.mark.timeout(600)
def test_s_per_query_beam_exact(oracle, cfg_s):
    from distributed_diskann_amd import datasets
    X, Q, w, idx, rows = cfg_s["X"], cfg_s["Q"], cfg_s["w"], cfg_s["idx"], cfg_s["rows"]
    beams = datasets.s_query_beams(Q.shape[0])
    _say("100K x 20: per-query-beam search of 10,000 queries")
    ids, dists, cmps, vis = idx.search(Q, 10, 200, 1.35, per_query_beam=beams)
    oi, od, oc, ov = oracle.search_all(oracle.Points(X, w.d), oracle.Graph(rows), Q, 10, 200, cut=1.35,
                                       per_query_beam=beams)
    assert np.array_equal(ids, oi)
    assert np.array_equal(dists.view(np.uint32), od.view(np.uint32))
    assert np.array_equal(cmps, oc) and np.array_equal(vis, ov)


# ---------------------------------------------------------------- 1M x 100 int8
@pytest.fixture(scope="module")
def cfg_i8(oracle):
    from distributed_diskann_amd import datasets
    X, Q, w = datasets.generate("i8")
    _say("1M x 100 int8: GPU build")
    idx, rows, secs = _build(X, w.d, w.metric, w.R, w.L, w.alpha)
    _say(f"1M x 100 int8: GPU build {secs:.2f} s")
    return dict(X=X, Q=Q, w=w, idx=idx, rows=rows)


@pytest.mark.timeout(900)
def test_i8_full_build_identical(oracle, cfg_i8):
    w = cfg_i8["w"]
    _say("1M x 100 int8: oracle build (about half a minute)")
    G = oracle.build(oracle.Points(cfg_i8["X"], w.d, w.metric), w.R, w.L, w.alpha, tie_mode=oracle.TIE_ID)
    assert np.array_equal(canon(cfg_i8["rows"]), canon(G.rows))


@pytest.mark.timeout(600)
@pytest.mark.parametrize("Ls", [40, 128])
def test_i8_search_exact(oracle, cfg_i8, Ls):
    _say(f"1M x 100 int8: search of 1,000 queries at Ls={Ls}")
    w = cfg_i8["w"]
    _search_exact(oracle, cfg_i8["idx"], cfg_i8["X"], cfg_i8["Q"][:1000], w.d, w.metric, cfg_i8["rows"], Ls)


# ---------------------------------------------------------------- 10M sets
def _big(name):
    from distributed_diskann_amd import datasets
    _say(f"{name}: generating the 10M-point set")
    X, Q, w = datasets.generate(name)
    _say(f"{name}: GPU build of 10M points")
    idx, rows, secs = _build(X, w.d, w.metric, w.R, w.L, w.alpha)
    _say(f"{name}: GPU build {secs:.2f} s")
    return dict(X=X, Q=Q, w=w, idx=idx, rows=rows)


@pytest.fixture(scope="module")
def cfg_u8(oracle):
    d = _big("u8")
    yield d
    d["idx"].close()


@pytest.fixture(scope="module")
def cfg_mips(oracle):
    d = _big("mips")
    yield d
    d["idx"].close()


def _prefix_identical(oracle, cfg, m=200_000):
    w = cfg["w"]
    Xp = np.ascontiguousarray(cfg["X"][:m])
    _say(f"{w.name}: GPU and oracle builds of the first {m} points")
    idx, rows, _ = _build(Xp, w.d, w.metric, w.R, w.L, w.alpha)
    idx.close()
    G = oracle.build(oracle.Points(Xp, w.d, w.metric), w.R, w.L, w.alpha, tie_mode=oracle.TIE_ID)
    assert np.array_equal(canon(rows), canon(G.rows))


def _rebuild_identical(cfg):
    w = cfg["w"]
    _say(f"{w.name}: invariants and a second 10M build")
    _invariants(cfg["rows"], w.R)
    idx, rows, _ = _build(cfg["X"], w.d, w.metric, w.R, w.L, w.alpha)
    idx.close()
    assert np.array_equal(rows, cfg["rows"])


@pytest.mark.timeout(900)
def test_u8_prefix_build_identical(oracle, cfg_u8):
    _prefix_identical(oracle, cfg_u8)


@pytest.mark.timeout(900)
def test_u8_full_build_invariants_and_rebuild(cfg_u8):
    _rebuild_identical(cfg_u8)


@pytest.mark.timeout(900)
def test_u8_search_exact_on_full_graph(oracle, cfg_u8):
    _say("u8: oracle search of 200 queries over the GPU-built 10M graph, Ls=140")
    w = cfg_u8["w"]
    _search_exact(oracle, cfg_u8["idx"], cfg_u8["X"], cfg_u8["Q"][:200], w.d, w.metric, cfg_u8["rows"], 140)


@pytest.mark.timeout(900)
def test_mips_prefix_build_identical(oracle, cfg_mips):
    _prefix_identical(oracle, cfg_mips)


@pytest.mark.timeout(900)
def test_mips_full_build_invariants_and_rebuild(cfg_mips):
    _rebuild_identical(cfg_mips)


@pytest.mark.timeout(900)
def test_mips_search_exact_on_full_graph(oracle, cfg_mips):
    _say("mips: oracle search of 200 queries over the GPU-built 10M graph, Ls=275")
    w = cfg_mips["w"]
    _search_exact(oracle, cfg_mips["idx"], cfg_mips["X"], cfg_mips["Q"][:200], w.d, w.metric, cfg_mips["rows"],
                  275)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("mode", ["bf16", "fp8"])
def test_mips_reduced_rerank_native_bits(oracle, cfg_mips, mode):
    _say(f"mips: {mode} traversal with FP32 re-rank of 100 over the 10M graph")
    X, Q, w, idx = cfg_mips["X"], cfg_mips["Q"][:500], cfg_mips["w"], cfg_mips["idx"]
    ids, dists, _, _ = idx.search(Q, 10, 250, 1.35, trav=mode, rerank=100)
    assert (ids < X.shape[0]).all()
    nat = oracle.distances("float32", "mips", np.repeat(Q, 10, axis=0), X[ids.reshape(-1)], w.d).reshape(ids.shape)
    assert np.array_equal(dists.view(np.uint32), nat.view(np.uint32))
    key = np.lexsort((ids, dists), axis=1)  # rows sorted by (distance, id)
    assert (key == np.arange(10)[None, :]).all()
