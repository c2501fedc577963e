"""Seeded synthetic datasets for the five benchmark workloads.

Every generator uses ``numpy.random.default_rng(seed)`` (PCG64) and draws in
fixed 1M-row chunks, so a dataset is a pure function of its seed and sizes.
Rows come back already padded to ``aligned_dims`` (point_range.h:45-50).

=====  ==========================================================  =========
name   base / queries                                             build
=====  ==========================================================  =========
xs     10,000 x 20 f32 N(0,1) seed 0; 1,000 queries seed 1          R32 L50
s      100,000 x 20 f32 N(0,1) seed 2; 10,000 queries seed 3        R32 L64
i8     1M x 100 int8, 1,000 centres N(0,30^2), noise N(0,12^2),     R64 L128
       rint, clip [-128,127], seed 4; queries seed 5
u8     10M x 128 uint8, 10,000 centres U[0,200], noise N(0,15^2),   R64 L128
       rint, clip [0,255], seed 6; queries seed 7
mips   10M x 200 f32, 5,000 centres N(0,1), scale 0.3, seed 8;      R64 L128
       queries from centres + N(0,0.5^2) offsets, seed 9            alpha 1.0
=====  ==========================================================  =========
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .formats import align_rows

CHUNK = 1 << 20


@dataclass(frozen=True)
class Workload:
    name: str
    n: int
    nq: int
    d: int
    dtype: str
    metric: str
    R: int
    L: int
    alpha: float
    base_seed: int
    query_seed: int


WORKLOADS = {
    "xs": Workload("xs", 10_000, 1_000, 20, "float32", "l2", 32, 50, 1.2, 0, 1),
    "s": Workload("s", 100_000, 10_000, 20, "float32", "l2", 32, 64, 1.2, 2, 3),
    "i8": Workload("i8", 1_000_000, 10_000, 100, "int8", "l2", 64, 128, 1.2, 4, 5),
    "u8": Workload("u8", 10_000_000, 10_000, 128, "uint8", "l2", 64, 128, 1.2, 6, 7),
    "mips": Workload("mips", 10_000_000, 10_000, 200, "float32", "mips", 64, 128, 1.0, 8, 9),
}

# per-query beam widths of the "s" workload: uniform over this set
S_BEAMS = np.array([10, 20, 50, 100, 200], np.uint32)
S_BEAM_SEED = 1003


def _gaussian(n, d, seed):
    rng = np.random.default_rng(seed)
    out = np.empty((n, d), np.float32)
    for s in range(0, n, CHUNK):
        e = min(n, s + CHUNK)
        out[s:e] = rng.standard_normal((e - s, d), dtype=np.float32)
    return out


def _mixture(n, d, centres, sigma, seed, lo, hi, dtype):
    rng = np.random.default_rng(seed)
    out = np.empty((n, d), dtype)
    k = centres.shape[0]
    for s in range(0, n, CHUNK):
        e = min(n, s + CHUNK)
        lab = rng.integers(0, k, e - s)
        x = centres[lab] + sigma * rng.standard_normal((e - s, d), dtype=np.float32)
        if lo is not None:
            x = np.clip(np.rint(x), lo, hi)
        out[s:e] = x.astype(dtype)
    return out


def generate(name: str, n: int | None = None, nq: int | None = None):
    """Returns ``(base_rows, query_rows, workload)`` (rows aligned).

    ``n``/``nq`` override the sizes (a prefix-compatible smaller instance of
    the same distribution, used by tests)."""
    w = WORKLOADS[name]
    n = w.n if n is None else n
    nq = w.nq if nq is None else nq
    if name in ("xs", "s"):
        X = _gaussian(n, w.d, w.base_seed)
        Q = _gaussian(nq, w.d, w.query_seed)
    elif name == "i8":
        crng = np.random.default_rng(w.base_seed)
        centres = (30.0 * crng.standard_normal((1000, w.d))).astype(np.float32)
        X = _mixture(n, w.d, centres, 12.0, w.base_seed + 1000, -128, 127, np.int8)
        Q = _mixture(nq, w.d, centres, 12.0, w.query_seed, -128, 127, np.int8)
    elif name == "u8":
        crng = np.random.default_rng(w.base_seed)
        centres = crng.uniform(0.0, 200.0, (10_000, w.d)).astype(np.float32)
        X = _mixture(n, w.d, centres, 15.0, w.base_seed + 1000, 0, 255, np.uint8)
        Q = _mixture(nq, w.d, centres, 15.0, w.query_seed, 0, 255, np.uint8)
    elif name == "mips":
        crng = np.random.default_rng(w.base_seed)
        centres = crng.standard_normal((5000, w.d)).astype(np.float32)
        X = _mixture(n, w.d, centres, 0.3, w.base_seed + 1000, None, None, np.float32)
        qrng = np.random.default_rng(w.query_seed)
        shifted = centres + (0.5 * qrng.standard_normal(centres.shape)).astype(np.float32)
        Q = _mixture(nq, w.d, shifted, 0.3, w.query_seed + 1000, None, None, np.float32)
    else:
        raise KeyError(name)
    return align_rows(X), align_rows(Q), w


def s_query_beams(nq: int = 10_000) -> np.ndarray:
    """Per-query beam widths of the "s" workload (seed 1003)."""
    rng = np.random.default_rng(S_BEAM_SEED)
    return S_BEAMS[rng.integers(0, len(S_BEAMS), nq)]
