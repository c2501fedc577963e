"""Reduced-precision traversal (BF16, per-dimension-scaled FP8-e4m3) with
FP32 re-rank, through the C-ABI.

* the device copies equal the numpy models of tests/lowp.py value for value
  (read back through one-hot MIPS queries, whose distance is -s_j * e_ij);
* traversal distances equal FP64 distances over the modelled copy to FP32
  accumulation accuracy, and stay within the copy's rounding bound of the
  FP64 distances over the original data;
* a re-ranked search returns exact FP32 distances (the NATIVE bits) sorted
  by (distance, id), and its recall stays within a stated margin of the
  FP32 traversal."""
import numpy as np
import pytest

from lowp import bf16_round, dist64, fp8_copy

pytestmark = pytest.mark.gpu


def _mips_data(n=20_000, nq=300):
    from distributed_diskann_amd import datasets
    X, Q, w = datasets.generate("mips", n=n, nq=nq)
    return X, Q, w.d


def _l2_data(n=4000, d=40, nq=64, seed=21):
    from distributed_diskann_amd.formats import align_rows
    rng = np.random.default_rng(seed)
    sc = rng.uniform(0.05, 4.0, d).astype(np.float32)
    X = (rng.standard_normal((n, d)).astype(np.float32) * sc)
    X[:, 3] = 0.0  # an all-zero dimension
    X[:7, 5] = np.float32(1e-3) * sc[5]  # values deep in the FP8 subnormal range of their dimension
    Q = (rng.standard_normal((nq, d)).astype(np.float32) * sc)
    return align_rows(X), align_rows(Q), d


def _onehot_read(X, d, mode, ids):
    """Decoded copy values -s_j * e_ij of rows `ids`, via one-hot MIPS queries."""
    from distributed_diskann_amd import Index
    idx = Index(X, d, "mips", 8)
    Q = np.zeros((d, X.shape[1]), np.float32)
    Q[np.arange(d), np.arange(d)] = 1.0
    out = idx.distances(Q, np.tile(ids[None, :], (d, 1)), mode)
    return -out.T  # (len(ids), d)


@pytest.mark.parametrize("data", ["mips", "l2"])
def test_bf16_copy_exact(data):
    print(f"[gpu] bf16 copy {data}", flush=True)
    X, _, d = _mips_data(3000, 1) if data == "mips" else _l2_data()
    ids = np.arange(0, X.shape[0], 7, dtype=np.uint32)[:400]
    got = _onehot_read(X, d, "bf16", ids)
    want = bf16_round(X[ids, :d])
    assert np.array_equal(got, want)


@pytest.mark.parametrize("data", ["mips", "l2"])
def test_fp8_copy_exact(data):
    print(f"[gpu] fp8 copy {data}", flush=True)
    X, _, d = _mips_data(3000, 1) if data == "mips" else _l2_data()
    ids = np.arange(0, X.shape[0], 7, dtype=np.uint32)[:400]
    ids[:7] = np.arange(7)
    got = _onehot_read(X, d, "fp8", ids)
    matches = []
    for variant in ("fn", "fnuz"):
        E, s = fp8_copy(X, d, variant)
        want = (s[None, :] * E[ids]).astype(np.float32)  # one IEEE float32 product, as the device forms it
        matches.append(np.array_equal(got, want))
    print(f"[gpu] fp8 copy matches OCP e4m3fn: {matches[0]}, FNUZ: {matches[1]}", flush=True)
    assert any(matches)


@pytest.mark.parametrize("metric", ["l2", "mips"])
@pytest.mark.parametrize("mode", ["bf16", "fp8"])
def test_traversal_distances(metric, mode):
    print(f"[gpu] traversal distances {metric} {mode}", flush=True)
    from distributed_diskann_amd import Index
    X, Q, d = _l2_data() if metric == "l2" else _mips_data(4000, 64)
    idx = Index(X, d, metric, 8)
    rng = np.random.default_rng(5)
    ids = rng.integers(0, X.shape[0], (Q.shape[0], 128)).astype(np.uint32)
    got = idx.distances(Q, ids, mode).astype(np.float64)
    if mode == "bf16":
        Xm = bf16_round(X[:, :d]).astype(np.float64)
        u = 2.0 ** -8
    else:
        E, s = fp8_copy(X, d, "fn")
        Xm = E.astype(np.float64) * s.astype(np.float64)
        u = 2.0 ** -4
    X64 = X[:, :d].astype(np.float64)
    for i in range(Q.shape[0]):
        q = Q[i, :d].astype(np.float64)
        model = dist64(Xm[ids[i]], q, metric)
        exact = dist64(X64[ids[i]], q, metric)
        # FP32 accumulation of d terms over the modelled copy
        mag = ((Xm[ids[i]] - q) ** 2).sum(1) if metric == "l2" else np.abs(Xm[ids[i]] * q).sum(1)
        assert np.all(np.abs(got[i] - model) <= 4e-6 * d ** 0.5 * mag + 1e-30)
        # rounding bound of the copy: |x - x^| <= u|x| (+ the FP8 subnormal half-quantum s_j 2^-10)
        dx = u * np.abs(X64[ids[i]]) + (s.astype(np.float64) * 2.0 ** -10 if mode == "fp8" else 0.0)
        if metric == "l2":
            bound = (dx * (2 * np.abs(X64[ids[i]] - q) + dx)).sum(1)
        else:
            bound = (dx * np.abs(q)).sum(1)
        assert np.all(np.abs(got[i] - exact) <= bound * (1 + 1e-5) + 4e-6 * d ** 0.5 * mag + 1e-30)


def _check_reranked(idx, Q, ids, dists, k):
    # exact FP32 distances of the returned ids, sorted by (distance, id), no duplicates
    nat = idx.distances(Q, ids, "native")
    assert np.array_equal(nat.view(np.uint32), dists.view(np.uint32))
    for i in range(ids.shape[0]):
        keys = list(zip(dists[i].tolist(), ids[i].tolist()))
        assert keys == sorted(keys)
        assert len(set(ids[i].tolist())) == k


def _recall(ids, gt_ids, gt_d, k=10):
    from distributed_diskann_amd.index import recall_with_ties
    return recall_with_ties(ids, gt_ids, gt_d, k)


def test_reduced_search_l2():
    print("[gpu] reduced search l2", flush=True)
    from distributed_diskann_amd import Index, datasets
    X, Q, w = datasets.generate("s", n=20_000, nq=500)
    idx = Index(X, w.d, "l2", 32)
    idx.build(R=32, L=64, alpha=1.2)
    gi, gd = idx.ground_truth(Q, 10)
    base, _, _, _ = idx.search(Q, 10, 40, 1.35)
    r0 = _recall(base, gi, gd)
    for mode, margin in (("bf16", 0.01), ("fp8", 0.02)):
        idx.prepare_traversal(mode)
        ids, dists, cmps, vis = idx.search(Q, 10, 40, 1.35, trav=mode, rerank=0)
        _check_reranked(idx, Q, ids, dists, 10)
        r = _recall(ids, gi, gd)
        print(f"[gpu] l2 recall native {r0:.4f} {mode} {r:.4f}", flush=True)
        assert r >= r0 - margin


def test_reduced_search_mips():
    print("[gpu] reduced search mips", flush=True)
    from distributed_diskann_amd import Index
    X, Q, d = _mips_data(20_000, 300)
    idx = Index(X, d, "mips", 32)
    idx.build(R=32, L=64, alpha=1.0)
    gi, gd = idx.ground_truth(Q, 10)
    base, _, _, _ = idx.search(Q, 10, 64, 1.35)
    r0 = _recall(base, gi, gd)
    for mode, margin in (("bf16", 0.01), ("fp8", 0.03)):
        ids, dists, _, _ = idx.search(Q, 10, 64, 1.35, trav=mode, rerank=0)
        _check_reranked(idx, Q, ids, dists, 10)
        r = _recall(ids, gi, gd)
        print(f"[gpu] mips recall native {r0:.4f} {mode} {r:.4f}", flush=True)
        assert r >= r0 - margin
        # re-ranking only the first 20 beam entries still returns exact, sorted distances
        ids2, d2, _, _ = idx.search(Q, 10, 64, 1.35, trav=mode, rerank=20)
        _check_reranked(idx, Q, ids2, d2, 10)


def test_reduced_search_device_and_beams():
    print("[gpu] reduced search device + per-query beams", flush=True)
    import torch
    from distributed_diskann_amd import Index, datasets
    X, Q, w = datasets.generate("s", n=8000, nq=256)
    idx = Index(X, w.d, "l2", 32)
    idx.build(R=32, L=64, alpha=1.2)
    beams = datasets.s_query_beams(256)
    hi, hd, hc, hv = idx.search(Q, 10, 200, 1.35, per_query_beam=beams, trav="fp8", rerank=100)
    dev = torch.device("cuda:0")
    Qd = torch.from_numpy(Q).to(dev)
    out_i = torch.empty((256, 10), dtype=torch.int32, device=dev)
    out_d = torch.empty((256, 10), dtype=torch.float32, device=dev)
    cm = torch.empty(256, dtype=torch.int64, device=dev)
    vs = torch.empty(256, dtype=torch.int32, device=dev)
    st = torch.zeros(1, dtype=torch.int32, device=dev)
    pqb = torch.from_numpy(beams.astype(np.int32)).to(dev)
    idx.search_device(Qd, 10, 200, out_i, out_d, cm, vs, cut=1.35, starts=st, per_query_beam=pqb, beam_max=200,
                      trav="fp8", rerank=100)
    torch.cuda.synchronize()
    assert np.array_equal(out_i.cpu().numpy().view(np.uint32), hi)
    assert np.array_equal(out_d.cpu().numpy().view(np.uint32), hd.view(np.uint32))
    assert np.array_equal(cm.cpu().numpy().view(np.uint64), hc)
    assert np.array_equal(vs.cpu().numpy().view(np.uint32), hv)


def test_reduced_errors():
    print("[gpu] reduced-precision argument errors", flush=True)
    from distributed_diskann_amd import Index
    from distributed_diskann_amd._lib import DDG_EINVAL, DdgError
    rng = np.random.default_rng(3)
    Xu = rng.integers(0, 255, (500, 64)).astype(np.uint8)
    iu = Index(Xu, 64, "l2", 8)
    with pytest.raises(DdgError) as e:
        iu.search(Xu[:4], 10, 20, 1.35, trav="bf16")
    assert e.value.status == DDG_EINVAL
    with pytest.raises(DdgError):
        iu.prepare_traversal("fp8")
    X, Q, d = _l2_data(500, 16, 4)
    idx = Index(X, d, "l2", 8)
    with pytest.raises(DdgError) as e:
        idx.search(Q, 10, 20, 1.35, trav="fp8", rerank=5)
    assert e.value.status == DDG_EINVAL


@pytest.mark.parametrize("case", ["uint8_100", "int8_100", "uint8_17", "int8_33"])
@pytest.mark.parametrize("qq", [0, 1])
def test_quantized_mips_golden(case, qq):
    """Quantized_Mips_Point::distance on the GPU equals the reference's own
    (golden vectors from its compiled mips_point.h) bit for bit."""
    print(f"[gpu] quantized inner products {case} quantized_query={qq}", flush=True)
    from conftest import load_golden
    from distributed_diskann_amd.index import quantized_mips_distances
    z = load_golden("qmips")
    got = quantized_mips_distances(z[case + "_codes"], z[f"{case}_q{qq}"], float(z[case + "_max"]),
                                   float(z[case + "_min"]))
    assert np.array_equal(got.view(np.uint32), z[f"{case}_out{qq}"].view(np.uint32))


@pytest.mark.parametrize("dtype", ["uint8", "int8"])
def test_quantized_mips_vs_oracle(oracle, dtype):
    print(f"[gpu] quantized inner products {dtype} 300 x 5000 d=128", flush=True)
    from distributed_diskann_amd.index import quantized_mips_distances
    rng = np.random.default_rng(41)
    lo, hi = (0, 256) if dtype == "uint8" else (-128, 128)
    codes = rng.integers(lo, hi, (5000, 128)).astype(dtype)
    for Q in (rng.standard_normal((300, 128)).astype(np.float32), rng.integers(lo, hi, (300, 128)).astype(dtype)):
        got = quantized_mips_distances(codes, Q, 0.75, -0.5)
        want = oracle.quantized_mips(codes, Q, 0.75, -0.5)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
