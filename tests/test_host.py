"""Host-side logic on the CPU: formats, parameters, datasets, and that the
C-ABI library loads and exports every symbol include/ddg.h declares (no
compute calls)."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

from conftest import ROOT

from distributed_diskann_amd import datasets, formats


def test_dim_round_up():
    # point_range.h:45-50 on the layouts the workloads use
    assert formats.dim_round_up(20, 4) == 32
    assert formats.dim_round_up(50, 4) == 64
    assert formats.dim_round_up(100, 1) == 128
    assert formats.dim_round_up(128, 1) == 128
    assert formats.dim_round_up(200, 4) == 208
    assert formats.dim_round_up(16, 4) == 16


def test_vector_and_graph_files_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    X = rng.standard_normal((37, 20)).astype(np.float32)
    p = str(tmp_path / "v")
    formats.write_vectors(p, X)
    Y, d = formats.read_vectors(p, "float32")
    assert d == 20 and Y.shape == (37, 32) and np.array_equal(Y[:, :20], X) and not Y[:, 20:].any()
    rows = np.zeros((37, 9), np.uint32)
    rows[:, 0] = rng.integers(0, 9, 37)
    for i in range(37):
        rows[i, 1:1 + rows[i, 0]] = rng.integers(0, 37, rows[i, 0])
    g = str(tmp_path / "g")
    formats.write_graph(g, rows)
    assert np.array_equal(formats.read_graph(g), rows)
    raw = np.fromfile(g, np.uint32)
    assert raw[0] == 37 and raw[1] == 8 and len(raw) == 2 + 37 + rows[:, 0].sum()


def test_truth_file_round_trip(tmp_path, xs):
    p = str(tmp_path / "gt")
    formats.write_truth(p, xs["gt_ids"], xs["gt_d"])
    i, d = formats.read_truth(p)
    assert np.array_equal(i, xs["gt_ids"]) and np.array_equal(d, xs["gt_d"])


def test_graph_with_degree_above_R_rejected(tmp_path):
    p = str(tmp_path / "g")
    with open(p, "wb") as f:
        np.array([2, 4, 5, 0], np.uint32).tofile(f)
        np.arange(5, dtype=np.uint32).tofile(f)
    with pytest.raises(ValueError):
        formats.read_graph(p)


def test_para_json(tmp_path):
    # the reference's para.json keys (dataset.hpp:64-89) and dataset JSON (92-125)
    para = {"build": True, "save": True, "query": False, "two_pass": False, "data_para": "./random-xs.json",
            "start_point_index": 0, "xita": 1, "alpha": 1.2, "L": 64, "R": 32, "Ls": 100,
            "num_neighbors_query": 10, "max": 0}
    ds = {"dataset": "random-xs", "nb": 10000, "nq": 1000, "d": 20, "dtype": "float32",
          "ds_fn": "data_10000_20", "qs_fn": "queries_1000_20", "gt_fn": "gt_10000_1000_20",
          "basedir": "data/random10000"}
    (tmp_path / "para.json").write_text(json.dumps(para))
    (tmp_path / "random-xs.json").write_text(json.dumps(ds))
    dp = formats.load_para(str(tmp_path / "para.json"), base_dir=str(tmp_path))
    assert (dp.R, dp.L, dp.alpha, dp.xita, dp.max_build) == (32, 64, 1.2, 1.0, 0)
    assert dp.ds_dir == "data/random10000/data_10000_20" and dp.output_path == "results/random-xs"
    del para["max"]
    (tmp_path / "bad.json").write_text(json.dumps(para))
    with pytest.raises(KeyError):
        formats.DatasetPara().set(str(tmp_path / "bad.json"))


def test_datasets_deterministic_and_shaped():
    X1, Q1, w = datasets.generate("xs", n=2000, nq=10)
    X2, Q2, _ = datasets.generate("xs", n=2000, nq=10)
    assert np.array_equal(X1, X2) and X1.shape == (2000, 32) and w.d == 20
    Xi, Qi, wi = datasets.generate("i8", n=3000, nq=5)
    assert Xi.dtype == np.int8 and Xi.shape == (3000, 128) and not Xi[:, 100:].any()
    Xu, Qu, wu = datasets.generate("u8", n=3000, nq=5)
    assert Xu.dtype == np.uint8 and Xu.shape == (3000, 128)
    Xm, Qm, wm = datasets.generate("mips", n=3000, nq=5)
    assert Xm.shape == (3000, 208) and wm.metric == "mips" and wm.alpha == 1.0
    b = datasets.s_query_beams(10000)
    assert set(np.unique(b)) == {10, 20, 50, 100, 200}


def test_recall_with_ties():
    from distributed_diskann_amd.index import recall_with_ties
    gt_i = np.array([[1, 2, 3, 4]], np.uint32)
    gt_d = np.array([[0.1, 0.2, 0.2, 0.3]], np.float32)
    # k=2: the GT id 3 ties the 2nd distance, so reporting {1, 3} scores 2/2
    assert recall_with_ties(np.array([[1, 3]]), gt_i, gt_d, 2) == 1.0
    assert recall_with_ties(np.array([[1, 4]]), gt_i, gt_d, 2) == 0.5


def _declared_symbols():
    txt = open(os.path.join(ROOT, "include", "ddg.h")).read()
    return sorted(set(re.findall(r"\b(ddg_[a-z_]+)\s*\(", txt)))


def test_cabi_exports_every_declared_symbol():
    from distributed_diskann_amd import _lib
    path = _lib.LIB_PATH
    if not os.path.exists(path):
        from distributed_diskann_amd import build
        build.build()
    so = ctypes.CDLL(path)  # loads without a GPU
    decl = _declared_symbols()
    assert len(decl) >= 15
    for name in decl:
        assert hasattr(so, name), name
    assert sorted(_lib.EXPORTS) == decl


def test_library_fails_loudly_without_so(monkeypatch, tmp_path):
    from distributed_diskann_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(ImportError):
        _lib.lib()


def test_create_without_gpu_reports_device_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from distributed_diskann_amd import DdgError, Index
    with pytest.raises(DdgError) as e:
        Index(np.zeros((4, 16), np.float32), 16, "l2", 4)
    assert e.value.status in (5,)


def test_lazy_delete_mirror():
    # index.h:139-164: ids go to delete_set, the start point is refused, an id
    # above G.size() is an error (G.size() itself passes, as in the reference)
    from distributed_diskann_amd import DdgError
    from distributed_diskann_amd.index import BuildParams, Graph, knn_index
    G = Graph(8, 100)
    I = knn_index(BuildParams(8, 16, 1.2, False))
    I.lazy_delete([3, 7, 0, 100], G)
    assert I.delete_set == {3, 7, 100}
    I.lazy_delete(5, G)
    assert 5 in I.delete_set
    with pytest.raises(DdgError):
        I.lazy_delete([101], G)
