"""Dense ground truth on matrix cores (FP64 MFMA for float32 rows, exact int8
MFMA for uint8/int8 rows) against the oracle's FP64 brute force.

Integer data: ids and distances bit-exact.  Float32 data: the GPU forms
|q|^2 + |x|^2 - 2 q.x in FP64 while the oracle sums (q - x)^2 directly, so
distances agree to FP64 rounding of the norms (relative 1e-12 of
|q|^2 + |x|^2) and ids agree except where the oracle's own neighbouring
distances tie to that tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _data(dtype, n, nq, d, seed):
    from distributed_diskann_amd.formats import align_rows
    rng = np.random.default_rng(seed)
    if dtype == "uint8":
        C = rng.uniform(0, 200, (30, d))
        X = np.clip(np.rint(C[rng.integers(0, 30, n)] + rng.normal(0, 15, (n, d))), 0, 255).astype(np.uint8)
        Q = np.clip(np.rint(C[rng.integers(0, 30, nq)] + rng.normal(0, 15, (nq, d))), 0, 255).astype(np.uint8)
    elif dtype == "int8":
        C = rng.normal(0, 30, (30, d))
        X = np.clip(np.rint(C[rng.integers(0, 30, n)] + rng.normal(0, 12, (n, d))), -128, 127).astype(np.int8)
        Q = np.clip(np.rint(C[rng.integers(0, 30, nq)] + rng.normal(0, 12, (nq, d))), -128, 127).astype(np.int8)
    else:
        X = rng.standard_normal((n, d)).astype(np.float32)
        Q = rng.standard_normal((nq, d)).astype(np.float32)
    return align_rows(X), align_rows(Q)


def gt_match(ids, d, oi, od, X, Q, dim, metric):
    """True when (ids, d) is the oracle's FP64 answer for float32 rows: every
    distance within 1e-12 (|q|^2 + |x|^2) of the oracle's, and an id that
    differs from the oracle's at some rank is a tie at that rank (its exact
    FP64 distance equals the oracle's there, to the same tolerance)."""
    Xd, Qd = X[:, :dim].astype(np.float64), Q[:, :dim].astype(np.float64)
    x2, q2 = (Xd ** 2).sum(1), (Qd ** 2).sum(1)
    tol = 1e-12 * (q2[:, None] + x2[oi])
    if not np.all(np.abs(d - od) <= tol):
        return False
    for i in range(ids.shape[0]):
        if len(set(ids[i].tolist())) != ids.shape[1]:
            return False
        for r in np.nonzero(ids[i] != oi[i])[0]:
            x = Xd[ids[i, r]]
            ex = ((x - Qd[i]) ** 2).sum() if metric == "l2" else -(x @ Qd[i])
            if abs(ex - od[i, r]) > 2 * tol[i, r]:
                return False
    return True


CASES = [
    # dtype, metric, n, nq, d, k
    ("float32", "l2", 20_000, 100, 200, 10),
    ("float32", "mips", 20_000, 70, 200, 10),
    ("float32", "l2", 5_003, 130, 20, 32),
    ("float32", "l2", 3_001, 65, 256, 1),
    ("float32", "mips", 2_000, 40, 300, 10),   # beyond the MFMA register budget: per-query kernel
    ("uint8", "l2", 30_011, 100, 128, 10),
    ("uint8", "mips", 10_000, 64, 100, 16),
    ("int8", "l2", 20_000, 100, 100, 10),
    ("int8", "mips", 7_777, 33, 48, 32),
    ("uint8", "l2", 999, 200, 512, 5),
    ("int8", "l2", 1_500, 10, 600, 10),        # beyond the int8 tile budget: per-query kernel
    # k = 100, the width of the reference's ground-truth files: longer
    # candidate lists (three entries per lane in a compaction), four-wave
    # int8 blocks
    ("float32", "mips", 30_000, 150, 200, 100),
    ("float32", "l2", 4_000, 70, 20, 100),
    ("uint8", "l2", 40_000, 150, 128, 100),
    ("int8", "l2", 20_000, 90, 100, 100),
    ("float32", "l2", 1_000, 9, 300, 100),     # per-query kernel
]


@pytest.mark.parametrize("dtype,metric,n,nq,d,k", CASES)
def test_ground_truth_mfma(oracle, dtype, metric, n, nq, d, k):
    print(f"[gpu] dense GT {dtype} {metric} n={n} nq={nq} d={d} k={k}", flush=True)
    from distributed_diskann_amd import Index
    X, Q = _data(dtype, n, nq, d, seed=n + d)
    idx = Index(X, d, metric, 8)
    ids, dd = idx.ground_truth(Q, k)
    oi, od = oracle.ground_truth(oracle.Points(X, d, metric), Q, k)
    if dtype == "float32":
        assert gt_match(ids, dd, oi, od, X, Q, d, metric)
    else:
        assert np.array_equal(ids, oi)
        assert np.array_equal(dd, od)


def test_ground_truth_duplicates_and_ties(oracle):
    print("[gpu] dense GT duplicate rows", flush=True)
    from distributed_diskann_amd import Index
    from distributed_diskann_amd.formats import align_rows
    rng = np.random.default_rng(9)
    base = rng.integers(0, 4, (50, 16)).astype(np.uint8)
    X = align_rows(base[rng.integers(0, 50, 4000)])  # many exact duplicates: ties broken by id
    Q = align_rows(base[rng.integers(0, 50, 77)])
    idx = Index(X, 16, "l2", 8)
    for k in (32, 100):
        ids, dd = idx.ground_truth(Q, k)
        oi, od = oracle.ground_truth(oracle.Points(X, 16, "l2"), Q, k)
        assert np.array_equal(ids, oi) and np.array_equal(dd, od)
