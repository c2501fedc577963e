"""The header-only C++ adapter (include/ddg/knn_index.hpp) used the way the
reference's own drivers call knn_index / searchAll (tests/cpp/adapter_demo.cpp):
it must compile and link against libddg.so on the CPU, and on the GPU give
the same graph and ids as the Python host mirror."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, canon

LIBDIR = os.path.join(ROOT, "distributed_diskann_amd")


def _compile(tmp_path):
    exe = str(tmp_path / "adapter_demo")
    cmd = ["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "adapter_demo.cpp"), "-L", LIBDIR, "-lddg", "-L", "/opt/rocm/lib",
           f"-Wl,-rpath,{LIBDIR}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def test_adapter_compiles_and_links(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libddg.so")):
        pytest.skip("libddg.so not built")
    assert os.path.exists(_compile(tmp_path))


@pytest.mark.gpu
def test_adapter_matches_python_mirror(tmp_path):
    print("[gpu] C++ adapter build + searchAll", flush=True)
    sys.path.insert(0, ROOT)
    from distributed_diskann_amd import Index, formats
    rng = np.random.default_rng(21)
    X = rng.integers(0, 256, (3000, 40), dtype=np.uint8)
    Q = rng.integers(0, 256, (50, 40), dtype=np.uint8)
    bf, qf = str(tmp_path / "base"), str(tmp_path / "q")
    formats.write_vectors(bf, X)
    formats.write_vectors(qf, Q)
    exe = _compile(tmp_path)
    gf, idf = str(tmp_path / "graph"), str(tmp_path / "ids")
    r = subprocess.run([exe, bf, qf, gf, idf, "24", "40", "1.2", "10", "50"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    G = formats.read_graph(gf)
    Xa, d = formats.read_vectors(bf, "uint8")
    idx = Index(Xa, d, "l2", 24)
    idx.build(R=24, L=40, alpha=1.2)
    assert np.array_equal(canon(G), canon(idx.get_graph()))
    ids = np.fromfile(idf, np.uint32).reshape(50, 10)
    Qa, _ = formats.read_vectors(qf, "uint8")
    want = idx.search(Qa, 10, 50, 1.35)[0]
    assert np.array_equal(ids, want)


DROPIN = os.path.join(ROOT, "oracle", "_ref", "dropin_demo")


@pytest.mark.gpu
@pytest.mark.timeout(600)
@pytest.mark.parametrize("Ls", [20, 100])
def test_dropin_on_reference_types(tmp_path, oracle, xs, Ls):
    """The adapter on the reference's own Graph<unsigned> / PointRange<float,
    Euclidian_Point<float>> (tests/cpp/dropin_demo.cpp, built from the
    reference's headers): beam_search (per point and batched) and searchAll
    equal the reference's own in the same process, bit for bit, also after a
    graph row is changed on the host; robustPrune of the reference's
    build-mode visited lists and build_index into a reference Graph equal the
    oracle's (distance, id)-ordered results."""
    if not os.path.exists(DROPIN):
        pytest.skip("oracle/_ref/dropin_demo is built only where the reference tree is present")
    print(f"[gpu] drop-in adapter on the reference's types, Ls={Ls}", flush=True)
    z = xs["z"]
    bf, qf, gf = tmp_path / "base", tmp_path / "queries", tmp_path / "graph"
    z["base_file"].tofile(bf)
    z["query_file"].tofile(qf)
    z["graph_file"].tofile(gf)
    out = str(tmp_path / "out")
    r = subprocess.run([DROPIN, str(bf), str(qf), str(gf), out, str(Ls), "32", "50", "1.2"], capture_output=True,
                       text=True, timeout=500)
    print(r.stdout[-2000:], flush=True)
    assert r.returncode == 0, (r.returncode, r.stdout[-3000:], r.stderr[-3000:])
    P = oracle.Points(xs["X"], xs["d"])
    G = oracle.Graph(xs["G"])
    raw = np.fromfile(out + ".prune", np.uint32)
    pos = 0
    for p in range(200):
        assert raw[pos] == p
        ln = int(raw[pos + 1])
        got = raw[pos + 2:pos + 2 + ln]
        pos += 2 + ln
        (_, _), (vi, vd), _ = oracle.beam_search(P, G, xs["X"][p], p, [0], 0, 50, 0.0, xs["X"].shape[0], 32)
        want = oracle.robust_prune(P, G, p, vi, vd, 1.2, add=True, R=32, tie_mode=oracle.TIE_ID)
        assert np.array_equal(got, want), p
    from distributed_diskann_amd import formats
    built = formats.read_graph(out + ".graph")
    want = oracle.build(P, 32, 50, 1.2, tie_mode=oracle.TIE_ID)
    assert np.array_equal(canon(built), canon(want.rows))
