"""The C-ABI's argument checks and layout rules on the GPU: every host scalar
of ddg_search_device is checked before a launch, a device start id outside
[0, n) is reported, a tuned launch shape that does not fit a later search
falls back to the default shape, and integer rows whose stride is not a
multiple of 64 bytes (the Gram-matrix prune's tile split) build exactly."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _xs_index(xs):
    from distributed_diskann_amd import Index
    idx = Index(xs["X"], xs["d"], "l2", xs["G"].shape[1] - 1)
    idx.set_graph(xs["G"])
    return idx


def test_search_device_rejects_bad_scalars(xs):
    print("[gpu] ddg_search_device argument checks", flush=True)
    import torch
    from distributed_diskann_amd._lib import DDG_EINVAL, DDG_ERANGE, QueryParamsC, lib
    idx = _xs_index(xs)
    dev = torch.device("cuda:0")
    nq = 32
    Q = torch.from_numpy(xs["Q"][:nq]).to(dev)
    ids = torch.empty((nq, 10), dtype=torch.int32, device=dev)
    starts = torch.zeros(80, dtype=torch.int32, device=dev)
    p = lambda t: C.c_void_p(t.data_ptr())

    def call(k=10, beam=50, cut=1.35, nst=1, beam_max=50, st=starts, trav=0, rerank=0):
        qp = QueryParamsC(k, beam, cut, 0, 0, trav, rerank)
        return lib().ddg_search_device(idx.h, p(Q), nq, None, p(st), nst, C.byref(qp), None, beam_max, p(ids),
                                       None, None, None)

    assert call() == 0
    assert call(nst=0) == DDG_EINVAL
    assert call(nst=65) == DDG_EINVAL  # only 64 start lanes exist
    assert call(beam_max=0) == DDG_EINVAL
    assert call(k=60, beam=50, beam_max=80) == DDG_EINVAL  # k > beam
    assert call(beam=90, beam_max=50) == DDG_EINVAL  # beam beyond the frontier capacity
    assert call(cut=0.5) == DDG_EINVAL  # cut < 1: the visited flag would not be exact
    assert call(trav=7) == DDG_EINVAL  # no such traversal mode
    assert call(trav=1, rerank=5) == DDG_EINVAL  # a re-rank shorter than k
    bad = torch.tensor([3, xs["X"].shape[0] + 5], dtype=torch.int32, device=dev)
    assert call(nst=2, st=bad) == DDG_ERANGE
    assert call() == 0  # the handle still works


def test_tuned_shape_falls_back_to_default(oracle, xs):
    print("[gpu] autotune at Ls=10, then search at Ls=64 with 64 starts", flush=True)
    idx = _xs_index(xs)
    idx.autotune_search(xs["Q"][:300], 10, 10)
    starts = list(range(0, 64 * 97, 97))
    ids, dists, cmps, vis = idx.search(xs["Q"], 10, 64, 1.35, starts=starts)
    P = oracle.Points(xs["X"], xs["d"])
    oi, od, oc, ov = oracle.search_all(P, oracle.Graph(xs["G"]), xs["Q"], 10, 64, cut=1.35, starts=starts)
    assert np.array_equal(ids, oi)
    assert np.array_equal(dists.view(np.uint32), od.view(np.uint32))
    assert np.array_equal(cmps, oc) and np.array_equal(vis, ov)


@pytest.mark.parametrize("dtype,d,ad", [("uint8", 100, 112), ("int8", 40, 48), ("uint8", 20, 32)])
def test_integer_rows_not_multiple_of_64_bytes(oracle, dtype, d, ad):
    print(f"[gpu] {dtype} d={d} rows of {ad} bytes", flush=True)
    from distributed_diskann_amd import Index
    rng = np.random.default_rng(d)
    lo, hi = (0, 256) if dtype == "uint8" else (-128, 128)
    X = np.zeros((3000, ad), dtype)
    X[:, :d] = rng.integers(lo, hi, (3000, d))
    P = oracle.Points(X, d)
    G = oracle.build(P, 24, 40, 1.2, tie_mode=oracle.TIE_ID)
    idx = Index(X, d, "l2", 24)
    idx.build(R=24, L=40, alpha=1.2)
    from conftest import canon
    assert np.array_equal(canon(idx.get_graph()), canon(G.rows))


def test_beam_search_reference_build_mode(xs):
    """ddg_beam_search in the build mode of index.h:277 (k = 0, L = 50,
    cut = 0) returns what the reference's own beam_search returned for these
    points (golden traces of the compiled reference): the whole frontier,
    the visited list in (distance, id) order and dist_cmps."""
    print("[gpu] beam_search build mode vs the reference's traces", flush=True)
    z = xs["z"]
    idx = _xs_index(xs)
    ids = z["ins_ids"]
    res = idx.beam_search(xs["X"][ids], k=0, beam=50, cut=0.0, limit=xs["X"].shape[0], degree_limit=32,
                          starts=(0,), query_ids=ids)
    vo = fo = 0
    for i, ((fi, fd), (vi, vd), c) in enumerate(res):
        vl, fl = int(z["ins_vlen"][i]), int(z["ins_fr_len"][i])
        assert c == int(z["ins_cmps"][i])
        assert np.array_equal(vi, z["ins_vis_ids"][vo:vo + vl])
        assert np.array_equal(vd.view(np.uint32), z["ins_vis_dists"][vo:vo + vl].view(np.uint32))
        assert np.array_equal(fi, z["ins_fr_ids"][fo:fo + fl])
        vo += vl
        fo += fl


@pytest.mark.parametrize("Ls,nstarts,limit", [(50, 1, 0), (120, 3, 0), (30, 1, 17), (10, 20, 0)])
def test_beam_search_query_mode_vs_oracle(oracle, xs, Ls, nstarts, limit):
    print(f"[gpu] beam_search query mode Ls={Ls} starts={nstarts} limit={limit}", flush=True)
    idx = _xs_index(xs)
    Q = xs["Q"][:200]
    starts = list(range(0, 97 * nstarts, 97))
    n = xs["X"].shape[0]
    res = idx.beam_search(Q, k=10, beam=Ls, cut=1.35, limit=limit, starts=starts)
    P = oracle.Points(xs["X"], xs["d"])
    G = oracle.Graph(xs["G"])
    for i, ((fi, fd), (vi, vd), c) in enumerate(res):
        (ofi, ofd), (ovi, ovd), oc = oracle.beam_search(P, G, Q[i], i, starts, 10, Ls, 1.35, limit or n, 32)
        assert c == oc
        assert np.array_equal(fi, ofi) and np.array_equal(fd.view(np.uint32), ofd.view(np.uint32))
        assert np.array_equal(vi, ovi) and np.array_equal(vd.view(np.uint32), ovd.view(np.uint32))


def test_beam_search_visited_longer_than_vcap(oracle, xs):
    print("[gpu] beam_search with a visited capacity below the lists", flush=True)
    from distributed_diskann_amd._lib import DDG_EINVAL
    idx = _xs_index(xs)
    short = idx.beam_search(xs["Q"][:20], k=10, beam=100, vcap=8)  # retried with the reported lengths
    full = idx.beam_search(xs["Q"][:20], k=10, beam=100)
    for a, b in zip(short, full):
        assert np.array_equal(a[1][0], b[1][0]) and a[2] == b[2]
    from distributed_diskann_amd._lib import QueryParamsC, lib
    Q = np.ascontiguousarray(xs["Q"][:4])
    vl = np.zeros(4, np.uint32)
    fi = np.zeros((4, 100), np.uint32)
    fd = np.zeros((4, 100), np.float32)
    fl = np.zeros(4, np.uint32)
    vi = np.zeros((4, 8), np.uint32)
    vd = np.zeros((4, 8), np.float32)
    st = np.zeros(1, np.uint32)
    qp = QueryParamsC(10, 100, 1.35, 0, 0, 0, 0)
    c = lambda a: C.c_void_p(a.ctypes.data)
    rc = lib().ddg_beam_search(idx.h, c(Q), 4, None, c(st), 1, C.byref(qp), 100, c(fi), c(fd), c(fl), 8, c(vi),
                               c(vd), c(vl), None)
    assert rc == DDG_EINVAL and (vl > 8).all()
    assert [int(v) for v in vl] == [len(r[1][0]) for r in full[:4]]


def test_graph_rows_round_trip(xs):
    print("[gpu] graph rows by id", flush=True)
    from distributed_diskann_amd import DdgError
    idx = _xs_index(xs)
    rng = np.random.default_rng(5)
    ids = rng.choice(xs["G"].shape[0], 300, replace=False)
    assert np.array_equal(idx.get_rows(ids), xs["G"][ids])
    new = xs["G"][ids[::-1]].copy()
    idx.set_rows(ids, new)
    full = idx.get_graph()
    want = xs["G"].copy()
    want[ids] = new
    assert np.array_equal(full, want)
    bad = new[:1].copy()
    bad[0, 0] = xs["G"].shape[1]  # degree above R
    with pytest.raises(DdgError):
        idx.set_rows(ids[:1], bad)
    bad = new[:1].copy()
    bad[0, 0], bad[0, 1] = 1, xs["G"].shape[0]  # neighbor id >= n
    with pytest.raises(DdgError):
        idx.set_rows(ids[:1], bad)
    assert np.array_equal(idx.get_graph(), want)  # rejected calls changed nothing
