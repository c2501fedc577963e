"""On-disk formats of the reference, read and written with numpy.

* vectors   ``[u32 n][u32 d][n*d T]``; rows are padded in memory to
  ``aligned_dims`` (point_range.h:45-50, 61-97)
* graph     ``[u32 n][u32 R][u32 deg[n]][u32 edges...]`` on disk,
  ``n x (R+1)`` rows ``[deg, nbr...]`` in memory (graph.h:126-200)
* truth     ``[u32 nq][u32 K][u32 ids nq*K][f32 dists nq*K]`` (types.h:42-90)
* params    ``para.json`` + dataset JSON (src/dataset.hpp:64-125)
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field

import numpy as np

NP_DTYPES = {"uint8": np.uint8, "int8": np.int8, "float32": np.float32}


def dim_round_up(dim: int, tp_size: int) -> int:
    """point_range.h:45-50: pad a row to a multiple of 64 bytes."""
    qt, rem = divmod(dim * tp_size, 64)
    return dim if rem == 0 else ((qt + 1) * 64) // tp_size


def align_rows(X: np.ndarray) -> np.ndarray:
    """Copy ``n x d`` rows into the zero-padded ``n x aligned_dims`` layout."""
    n, d = X.shape
    ad = dim_round_up(d, X.dtype.itemsize)
    if ad == d:
        return np.ascontiguousarray(X)
    out = np.zeros((n, ad), X.dtype)
    out[:, :d] = X
    return out


def read_vectors(path: str, dtype: str, aligned: bool = True):
    """Returns ``(rows, d)``; rows are aligned unless ``aligned`` is False."""
    with open(path, "rb") as f:
        n, d = np.fromfile(f, np.uint32, 2).tolist()
        X = np.fromfile(f, NP_DTYPES[dtype], n * d).reshape(n, d)
    return (align_rows(X) if aligned else X), d


def write_vectors(path: str, X: np.ndarray, d: int | None = None):
    d = X.shape[1] if d is None else d
    with open(path, "wb") as f:
        np.asarray([X.shape[0], d], np.uint32).tofile(f)
        np.ascontiguousarray(X[:, :d]).tofile(f)


def read_graph(path: str) -> np.ndarray:
    """graph.h:130-176."""
    with open(path, "rb") as f:
        n, R = np.fromfile(f, np.uint32, 2).tolist()
        deg = np.fromfile(f, np.uint32, n)
        edges = np.fromfile(f, np.uint32, int(deg.sum()))
    if np.any(deg > R):
        raise ValueError(f"{path}: a degree exceeds R={R}")
    rows = np.zeros((n, R + 1), np.uint32)
    rows[:, 0] = deg
    mask = np.arange(R)[None, :] < deg[:, None]
    rows[:, 1:][mask] = edges
    return rows


def write_graph(path: str, rows: np.ndarray):
    """graph.h:178-200."""
    n, R1 = rows.shape
    deg = rows[:, 0]
    mask = np.arange(R1 - 1)[None, :] < deg[:, None]
    with open(path, "wb") as f:
        np.asarray([n, R1 - 1], np.uint32).tofile(f)
        deg.astype(np.uint32).tofile(f)
        rows[:, 1:][mask].astype(np.uint32).tofile(f)


def read_truth(path: str):
    """types.h:42-66: ids (u32) and distances (f32)."""
    with open(path, "rb") as f:
        nq, K = np.fromfile(f, np.uint32, 2).tolist()
        ids = np.fromfile(f, np.uint32, nq * K).reshape(nq, K)
        dists = np.fromfile(f, np.float32, nq * K).reshape(nq, K)
    return ids, dists


def write_truth(path: str, ids: np.ndarray, dists: np.ndarray):
    with open(path, "wb") as f:
        np.asarray(ids.shape, np.uint32).tofile(f)
        ids.astype(np.uint32).tofile(f)
        dists.astype(np.float32).tofile(f)


@dataclass
class DatasetPara:
    """src/dataset.hpp:30-134: para.json keys plus the dataset JSON keys."""
    build: bool = True
    query: bool = True
    save: bool = True
    two_pass: bool = False
    start_point_index: int = 0
    xita: float = 0.02
    alpha: float = 1.2
    L: int = 128
    R: int = 64
    Ls: int = 128
    num_neighbors_query: int = 100
    max_build: int = 0
    dataset_parafile: str = ""
    dataset_name: str = ""
    nb: int = 0
    nq: int = 0
    d: int = 0
    dtype: str = ""
    ds_fn: str = ""
    qs_fn: str = ""
    gt_fn: str = ""
    basedir: str = ""
    device: dict = field(default_factory=dict)

    @property
    def ds_dir(self):
        return self.basedir + "/" + self.ds_fn

    @property
    def qs_dir(self):
        return self.basedir + "/" + self.qs_fn

    @property
    def gt_dir(self):
        return self.basedir + "/" + self.gt_fn

    @property
    def output_path(self):
        return "results/" + self.dataset_name

    def set(self, filename: str):
        """dataset.hpp:64-89 (every key is required, as in the reference)."""
        with open(filename) as f:
            j = json.load(f)
        self.build = bool(j["build"]); self.query = bool(j["query"]); self.save = bool(j["save"])
        self.two_pass = bool(j["two_pass"]); self.start_point_index = int(j["start_point_index"])
        self.xita = float(j["xita"]); self.alpha = float(j["alpha"]); self.L = int(j["L"]); self.R = int(j["R"])
        self.Ls = int(j["Ls"]); self.max_build = int(j["max"])
        self.num_neighbors_query = int(j["num_neighbors_query"])
        self.dataset_parafile = str(j["data_para"])
        self.device = dict(j.get("device", {}))  # optional device-only block
        return self

    def open(self, filename: str):
        """dataset.hpp:92-125."""
        with open(filename) as f:
            j = json.load(f)
        self.dataset_name = j["dataset"]; self.nb = int(j["nb"]); self.nq = int(j["nq"]); self.d = int(j["d"])
        self.dtype = j["dtype"]; self.ds_fn = j["ds_fn"]; self.qs_fn = j["qs_fn"]; self.gt_fn = j["gt_fn"]
        self.basedir = j["basedir"]
        return self


def load_para(para_file: str, base_dir: str | None = None) -> DatasetPara:
    dp = DatasetPara().set(para_file)
    ds = dp.dataset_parafile
    if base_dir is not None and not os.path.isabs(ds):
        ds = os.path.join(base_dir, ds)
    return dp.open(ds)
