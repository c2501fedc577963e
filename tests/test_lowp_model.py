"""CPU pins of the numpy models in tests/lowp.py: closed-form rounding cases
and torch's own float8 / bfloat16 conversions (round to nearest even)."""
import numpy as np
import pytest
import torch

from lowp import bf16_round, e4m3_round, fp8_copy


def test_bf16_closed_form():
    x = np.array([1.0, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, 1 + 2.0 ** -8 + 2.0 ** -20, -2.5, 3.0e38, 0.0],
                 np.float32)
    want = np.array([1.0, 1.0, 1 + 2.0 ** -6, 1 + 2.0 ** -7, -2.5, np.float32(3.0e38), 0.0], np.float64)
    got = bf16_round(x).astype(np.float64)
    # 3.0e38 rounds to the nearest bf16 value, whatever it is; check the rest exactly
    assert np.array_equal(got[[0, 1, 2, 3, 4, 6]], want[[0, 1, 2, 3, 4, 6]])
    assert abs(got[5] - 3.0e38) / 3.0e38 <= 2.0 ** -8


def test_bf16_matches_torch():
    rng = np.random.default_rng(11)
    x = (rng.standard_normal(100_000) * np.exp(rng.uniform(-20, 20, 100_000))).astype(np.float32)
    t = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(bf16_round(x), t)


@pytest.mark.parametrize("variant,cases", [
    ("fn", [(448.0, 448.0), (240.0, 240.0), (1.0625, 1.0), (1.1875, 1.25), (17.0, 16.0), (2.0 ** -9, 2.0 ** -9),
            (2.0 ** -10, 0.0), (3 * 2.0 ** -10, 2.0 ** -8), (2.0 ** -6, 2.0 ** -6), (-0.3, -0.3125)]),
    ("fnuz", [(240.0, 240.0), (1.0625, 1.0), (2.0 ** -10, 2.0 ** -10), (2.0 ** -11, 0.0),
              (3 * 2.0 ** -11, 2.0 ** -9), (2.0 ** -7, 2.0 ** -7)]),
])
def test_e4m3_closed_form(variant, cases):
    v = np.array([c[0] for c in cases], np.float64)
    want = np.array([c[1] for c in cases], np.float32)
    assert np.array_equal(e4m3_round(v, variant), want)


@pytest.mark.parametrize("variant,tdtype", [("fn", torch.float8_e4m3fn), ("fnuz", torch.float8_e4m3fnuz)])
def test_e4m3_matches_torch(variant, tdtype):
    rng = np.random.default_rng(12)
    x = rng.uniform(-240, 240, 200_000).astype(np.float32)
    x[:50_000] *= np.float32(2.0 ** -12)  # subnormal range
    t = torch.from_numpy(x).to(tdtype).to(torch.float32).numpy()
    assert np.array_equal(e4m3_round(x, variant), t)


def test_fp8_copy_range_and_scale():
    rng = np.random.default_rng(13)
    X = (rng.standard_normal((1000, 24)) * rng.uniform(0.01, 50, 24)).astype(np.float32)
    X[:, 5] = 0.0  # an all-zero dimension gets scale 1
    E, s = fp8_copy(X, 24)
    assert np.abs(E).max() <= 240.0
    assert s[5] == 1.0 and np.all(E[:, 5] == 0)
    # every dimension's largest magnitude maps to +-240 exactly
    j = [c for c in range(24) if c != 5]
    assert np.allclose(np.abs(E[:, j]).max(0), 240.0)
    # the decoded copy is within half an e4m3 step (2^-4 relative) of the data, or the subnormal half-quantum
    err = np.abs(E.astype(np.float64) * s - X)
    assert np.all(err <= np.abs(X) * 2.0 ** -4 + s * 2.0 ** -10 + 1e-30)
