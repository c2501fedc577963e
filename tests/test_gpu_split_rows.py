"""Native float32 rows split over lanes (GPU).

A float32 row of more than eight 16-byte chunks is spread over S lanes (2, 4,
8, 16, ...: the fewest leaving a lane at most 8 chunks, each instantiation
walking exactly the lane's 5-8 chunks; rows of more than 64 x 8 chunks keep
one lane per row), which load their
chunks together and hand the running FP32 sum along in the reference's
order (euclidian_point.h / NSGDist.h:113-131 for L2, mips_point.h:60-66 for
MIPS).  Builds must equal the oracle's row for row and searches must be
bit-exact (ids, distance bits, dist_cmps, |visited|) for every split width,
for both metrics, and for the L2 rows with d % 4 != 0 that keep one lane per
row; the same searches with the split turned off (DDG_NAT_SPLIT=0) must
return the same bits."""
import numpy as np
import pytest

from conftest import canon

pytestmark = pytest.mark.gpu

# (metric, d): aligned row chunks 12 (S=2), 28 (S=4), 52 (S=8), 76 (S=16); d = 37 (L2) keeps one lane per row,
# and so does d = 2100 (528 chunks: more than 64 lanes of 8)
CASES = [("l2", 40), ("mips", 100), ("l2", 200), ("mips", 300), ("l2", 300), ("l2", 37), ("mips", 2100)]


def _data(d, n=3000, nq=200, seed=3):
    from distributed_diskann_amd.formats import align_rows
    rng = np.random.default_rng(seed + d)
    C = rng.standard_normal((30, d)).astype(np.float32)
    X = (C[rng.integers(0, 30, n)] + 0.4 * rng.standard_normal((n, d))).astype(np.float32)
    Q = (C[rng.integers(0, 30, nq)] + 0.4 * rng.standard_normal((nq, d))).astype(np.float32)
    return align_rows(X), align_rows(Q)


@pytest.mark.parametrize("metric,d", CASES)
def test_split_rows_match_oracle(oracle, monkeypatch, metric, d):
    from distributed_diskann_amd import Index
    print(f"[gpu] split rows {metric} d={d}", flush=True)
    X, Q = _data(d)
    alpha = 1.0 if metric == "mips" else 1.2
    P = oracle.Points(X, d, metric)
    G = oracle.build(P, 24, 48, alpha, tie_mode=oracle.TIE_ID)
    idx = Index(X, d, metric, 24)
    idx.build(R=24, L=48, alpha=alpha)
    rows = idx.get_graph()
    assert np.array_equal(canon(rows), canon(G.rows))
    for Ls in (30, 120):
        ids, dists, cmps, vis = idx.search(Q, 10, Ls, 1.35)
        oi, od, oc, ov = oracle.search_all(P, oracle.Graph(rows), Q, 10, Ls, cut=1.35)
        assert np.array_equal(ids, oi), f"Ls={Ls}"
        assert np.array_equal(dists.view(np.uint32), od.view(np.uint32)), f"Ls={Ls}"
        assert np.array_equal(cmps, oc) and np.array_equal(vis, ov), f"Ls={Ls}"
        monkeypatch.setenv("DDG_NAT_SPLIT", "0")
        one = idx.search(Q, 10, Ls, 1.35)
        monkeypatch.delenv("DDG_NAT_SPLIT")
        assert all(np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a, b.view(np.uint32)
                                  if b.dtype == np.float32 else b) for a, b in zip(one, (ids, dists, cmps, vis)))
    idx.close()
