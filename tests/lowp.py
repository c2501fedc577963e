"""numpy models of the reduced traversal copies (test infrastructure).

bf16_round    float32 -> nearest bfloat16 (ties to even), as float32
e4m3_round    float -> nearest FP8-e4m3 value (ties to even) in the OCP
              ("fn": bias 7, subnormal quantum 2^-9) or FNUZ (bias 8,
              subnormal quantum 2^-10) variant; both agree on normals and
              differ only below 2^-6
fp8_copy      the per-dimension scaled copy: s_j = max_i |x_ij| / 240 and
              codes e_ij = e4m3(x_ij * (240 / max_j)) in IEEE float32
"""
from __future__ import annotations

import numpy as np


def bf16_round(x: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def e4m3_round(v: np.ndarray, variant: str = "fn") -> np.ndarray:
    bias = {"fn": 7, "fnuz": 8}[variant]
    emin = 1 - bias
    a = np.abs(np.asarray(v, np.float64))
    _, ex = np.frexp(a)  # a = m * 2^ex, m in [0.5, 1)
    e = np.maximum(ex - 1, emin)
    q = np.ldexp(1.0, e - 3)
    r = np.round(a / q) * q  # np.round: ties to even
    return (np.sign(v) * r).astype(np.float32)


def fp8_scales(X: np.ndarray, d: int):
    m = np.abs(np.asarray(X, np.float32)[:, :d]).max(0)
    ok = np.isfinite(m) & (m >= np.float32(1e-30))
    s = np.where(ok, m / np.float32(240.0), np.float32(1.0)).astype(np.float32)
    inv = np.where(ok, np.float32(240.0) / np.where(ok, m, 1), np.float32(1.0)).astype(np.float32)
    return s, inv


def fp8_copy(X: np.ndarray, d: int, variant: str = "fn"):
    """(codes as float32 values, s) of the FP8 traversal copy."""
    s, inv = fp8_scales(X, d)
    v = (np.asarray(X, np.float32)[:, :d] * inv[None, :]).astype(np.float32)
    return e4m3_round(v, variant), s


def dist64(Xm: np.ndarray, q: np.ndarray, metric: str) -> np.ndarray:
    """FP64 distances of the rows Xm (n x d) to q (d)."""
    Xm = np.asarray(Xm, np.float64)
    q = np.asarray(q, np.float64)
    if metric == "l2":
        return ((Xm - q[None, :]) ** 2).sum(1)
    return -(Xm @ q)
