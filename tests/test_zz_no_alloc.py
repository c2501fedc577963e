"""ddg_search_device_noalloc (no device allocation; status words and
visited-filter slices in the library's fixed per-GPU scratch array) against
ddg_search_device through the C-ABI: identical ids, distance bits, dist_cmps
and |visited| on two calls in a row, the second with fewer queries, for a
beam whose filter sits in LDS, one whose filter slices sit in the scratch
array, per-query beams and a reduced-precision traversal."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _index(xs):
    from distributed_diskann_amd import Index
    idx = Index(xs["X"], xs["d"], "l2", xs["G"].shape[1] - 1)
    idx.set_graph(xs["G"])
    return idx


class Outs:
    def __init__(self, torch, nq, k):
        dev = torch.device("cuda:0")
        self.ids = torch.full((nq, k), -7, dtype=torch.int32, device=dev)
        self.d = torch.zeros((nq, k), dtype=torch.float32, device=dev)
        self.cm = torch.zeros(nq, dtype=torch.int64, device=dev)
        self.vs = torch.zeros(nq, dtype=torch.int32, device=dev)

    def host(self):
        return (self.ids.cpu().numpy(), self.d.cpu().numpy().view(np.uint32), self.cm.cpu().numpy(),
                self.vs.cpu().numpy())


def _run(fn, idx, Q, nq, o, starts, k, beam, beam_max, pqb=None, trav=0, rerank=0):
    from distributed_diskann_amd._lib import QueryParamsC
    qp = QueryParamsC(k, beam, 1.35, 0, 0, trav, rerank)
    p = lambda t: None if t is None else C.c_void_p(t.data_ptr())
    return fn(idx.h, p(Q), nq, None, p(starts), starts.numel(), C.byref(qp), p(pqb), beam_max, p(o.ids), p(o.d),
              p(o.cm), p(o.vs))


@pytest.mark.parametrize("case", ["lds_filter", "scratch_filter", "per_query_beams", "bf16"])
def test_noalloc_matches_search_device(xs, case):
    print(f"[gpu] allocation-free search vs ddg_search_device: {case}", flush=True)
    import torch
    from distributed_diskann_amd._lib import lib
    idx = _index(xs)
    dev = torch.device("cuda:0")
    Qall = torch.from_numpy(xs["Q"]).to(dev)
    starts = torch.zeros(1, dtype=torch.int32, device=dev)
    k, beam, trav, rerank, pqb = 10, 50, 0, 0, None
    if case == "scratch_filter":
        beam = 160  # 2^13-slot filters in global memory
    if case == "per_query_beams":
        rng = np.random.default_rng(3)
        beam = 200
        pqb = torch.from_numpy(np.array([10, 20, 50, 100, 200], np.int32)[rng.integers(0, 5, 1000)]).to(dev)
    if case == "bf16":
        trav, rerank = 1, 20
        idx.prepare_traversal("bf16")
    for nq in (1000, 300):  # the second call has fewer queries
        Q = Qall[:nq].contiguous()
        a, b = Outs(torch, nq, k), Outs(torch, nq, k)
        pq = None if pqb is None else pqb[:nq].contiguous()
        assert _run(lib().ddg_search_device, idx, Q, nq, a, starts, k, beam, beam, pq, trav, rerank) == 0
        assert _run(lib().ddg_search_device_noalloc, idx, Q, nq, b, starts, k, beam, beam, pq, trav, rerank) == 0
        for x, y in zip(a.host(), b.host()):
            assert np.array_equal(x, y)
        assert (a.host()[0] >= 0).all()


def test_noalloc_refusals(xs):
    print("[gpu] allocation-free search refusals", flush=True)
    import torch
    from distributed_diskann_amd._lib import DDG_EINVAL, DDG_ERANGE, lib
    idx = _index(xs)
    dev = torch.device("cuda:0")
    Q = torch.from_numpy(xs["Q"][:64]).to(dev)
    o = Outs(torch, 64, 10)
    starts = torch.zeros(1, dtype=torch.int32, device=dev)
    fn = lib().ddg_search_device_noalloc
    # the FP8 copy was never prepared on this handle: no allocation allowed
    assert _run(fn, idx, Q, 64, o, starts, 10, 50, 50, trav=2, rerank=20) == DDG_EINVAL
    # beam 8192: one wave per block and a 2^24-slot (64 MiB) filter, beyond the scratch
    assert _run(fn, idx, Q, 64, o, starts, 10, 8192, 8192) == DDG_EINVAL
    bad = torch.tensor([xs["X"].shape[0]], dtype=torch.int32, device=dev)
    assert _run(fn, idx, Q, 64, o, bad, 10, 50, 50) == DDG_ERANGE
    assert _run(fn, idx, Q, 64, o, starts, 10, 50, 50) == 0


def test_noalloc_leaves_handle_unallocated(xs):
    print("[gpu] allocation-free search on a fresh handle allocates no buffer", flush=True)
    import torch
    from distributed_diskann_amd._lib import lib
    L = lib()
    L.ddg_internal_buffer_bytes.restype = C.c_uint64
    L.ddg_internal_buffer_bytes.argtypes = [C.c_void_p]
    idx = _index(xs)  # ddg_create + ddg_graph_set: no scratch buffer yet
    assert L.ddg_internal_buffer_bytes(idx.h) == 0
    dev = torch.device("cuda:0")
    starts = torch.zeros(1, dtype=torch.int32, device=dev)
    for beam, nq in ((160, 500), (50, 200)):
        Q = torch.from_numpy(xs["Q"][:nq]).to(dev)
        o = Outs(torch, nq, 10)
        assert _run(L.ddg_search_device_noalloc, idx, Q, nq, o, starts, 10, beam, beam) == 0
        assert (o.host()[0] >= 0).all()
    assert L.ddg_internal_buffer_bytes(idx.h) == 0
