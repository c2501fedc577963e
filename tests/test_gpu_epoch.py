"""Epoch-tagged global visited filter (GPU).

Beams whose filter is above 8 KB keep it in global memory, one slice per
resident wave.  The slices are epoch-tagged: a stored word is
id | epoch << id_bits and each query advances its slice's epoch, so no
query clears its slice; the slice is cleared only when the epoch wraps, and
the pool is set up again whenever its slice size changes or another use
wrote untagged ids into it.  None of this may change a result:

* searches at beams with different filter sizes, in an order that switches
  the pool's layout back and forth and after a build has used the pool,
  equal the oracle bit for bit (native traversal) and equal the same
  searches with per-query clearing (DDG_EPOCH=0; all traversal modes);
* with the epoch cycle cut to 2 and 3 (DDG_EPOCH_END), so that slices wrap
  and are cleared every one or two queries across launches, the results are
  the same again."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BEAMS = [120, 250, 120, 300, 250]  # 2^12, 2^14, 2^12, 2^15, 2^14 filter slots


def _say(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def mips():
    from distributed_diskann_amd import Index, datasets
    X, Q, w = datasets.generate("mips", n=30_000, nq=400)
    idx = Index(X, w.d, w.metric, 32)
    idx.build(R=32, L=64, alpha=1.0)
    rows = idx.get_graph()
    idx.prepare_traversal("bf16")
    idx.prepare_traversal("fp8")
    yield X, Q, w.d, idx, rows
    idx.close()


def _run(idx, Q, mode, Ls):
    rr = 0 if mode == "native" else 100
    ids, dists, cmps, vis = idx.search(Q, 10, Ls, 1.35, trav=mode, rerank=rr)
    return ids, dists.view(np.uint32), cmps, vis


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def _with_env(name, value, fn):
    old = os.environ.get(name)
    os.environ[name] = value
    try:
        return fn()
    finally:
        if old is None:
            del os.environ[name]
        else:
            os.environ[name] = old


def test_epoch_native_matches_oracle(mips, oracle):
    X, Q, d, idx, rows = mips
    P, G = oracle.Points(X, d, "mips"), oracle.Graph(rows)
    for Ls in BEAMS:
        _say(f"epoch native Ls={Ls} vs oracle")
        got = _run(idx, Q, "native", Ls)
        oi, od, oc, ov = oracle.search_all(P, G, Q, 10, Ls, cut=1.35)
        assert _same(got, (oi, od.view(np.uint32), oc, ov)), f"Ls={Ls}"


@pytest.mark.parametrize("mode", ["native", "bf16", "fp8"])
def test_epoch_equals_clearing(mips, mode):
    X, Q, d, idx, rows = mips
    clear = {Ls: _with_env("DDG_EPOCH", "0", lambda: _run(idx, Q, mode, Ls)) for Ls in sorted(set(BEAMS))}
    for Ls in BEAMS:
        _say(f"epoch {mode} Ls={Ls} vs per-query clearing")
        assert _same(_run(idx, Q, mode, Ls), clear[Ls]), f"{mode} Ls={Ls}"
    for end in ("2", "3"):
        for Ls in BEAMS:
            _say(f"epoch {mode} Ls={Ls} with the epoch cycle cut to {end}")
            for _ in range(2):  # wraps carry across launches
                got = _with_env("DDG_EPOCH_END", end, lambda: _run(idx, Q, mode, Ls))
                assert _same(got, clear[Ls]), f"{mode} Ls={Ls} epoch_end={end}"


def test_epoch_after_build_reuses_pool(oracle):
    """A build's insert searches (global filter of another slice size, or the
    written-slot bitmap) use the same pool; the next search must not read
    their words as tagged ids."""
    from distributed_diskann_amd import Index, datasets
    X, Q, w = datasets.generate("mips", n=12_000, nq=200)
    idx = Index(X, w.d, w.metric, 32)
    _say("search, rebuild, search again")
    idx.build(R=32, L=128, alpha=1.0)
    rows = idx.get_graph()
    first = _run(idx, Q, "native", 250)
    idx.set_graph(np.zeros_like(rows))  # build_index inserts into the graph it is given: start empty again
    idx.build(R=32, L=128, alpha=1.0)
    assert np.array_equal(idx.get_graph(), rows)
    again = _run(idx, Q, "native", 250)
    oi, od, oc, ov = oracle.search_all(oracle.Points(X, w.d, "mips"), oracle.Graph(rows), Q, 10, 250, cut=1.35)
    want = (oi, od.view(np.uint32), oc, ov)
    assert _same(first, want) and _same(again, want)
    idx.close()
