import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libddg.so)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def parse_vectors(buf: np.ndarray, dtype: str):
    """Reference vector file bytes -> (aligned rows, d)."""
    from distributed_diskann_amd import formats
    hdr = buf[:8].view(np.uint32)
    n, d = int(hdr[0]), int(hdr[1])
    X = buf[8:8 + n * d * np.dtype(dtype).itemsize].view(dtype).reshape(n, d)
    return formats.align_rows(X), d


def parse_graph(buf: np.ndarray):
    hdr = buf[:8].view(np.uint32)
    n, R = int(hdr[0]), int(hdr[1])
    deg = buf[8:8 + 4 * n].view(np.uint32)
    edges = buf[8 + 4 * n:].view(np.uint32)
    rows = np.zeros((n, R + 1), np.uint32)
    rows[:, 0] = deg
    mask = np.arange(R)[None, :] < deg[:, None]
    rows[:, 1:][mask] = edges
    return rows


def parse_truth(buf: np.ndarray):
    hdr = buf[:8].view(np.uint32)
    nq, K = int(hdr[0]), int(hdr[1])
    ids = buf[8:8 + 4 * nq * K].view(np.uint32).reshape(nq, K)
    d = buf[8 + 4 * nq * K:].view(np.float32).reshape(nq, K)
    return ids, d


@pytest.fixture(scope="session")
def xs():
    z = load_golden("xs")
    X, d = parse_vectors(z["base_file"], "float32")
    Q, _ = parse_vectors(z["query_file"], "float32")
    gi, gd = parse_truth(z["gt_file"])
    G = parse_graph(z["graph_file"])
    return dict(z=z, X=X, Q=Q, d=d, gt_ids=gi, gt_d=gd, G=G)


@pytest.fixture(scope="session")
def n0():
    from distributed_diskann_amd import datasets
    z = load_golden("n0")
    X, Q, w = datasets.generate("xs")
    G = parse_graph(z["graph_file"])
    return dict(z=z, X=X, Q=Q, d=w.d, G=G, w=w)


def canon(rows: np.ndarray) -> np.ndarray:
    """Graph rows with everything past each row's degree zeroed (the part
    Graph::save keeps)."""
    out = rows.copy()
    mask = np.arange(rows.shape[1] - 1)[None, :] >= rows[:, :1]
    out[:, 1:][mask] = 0
    return out


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
