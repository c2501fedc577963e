"""The recall / QPS reporting protocol (check_nn_recall.h, parse_results.h,
stats.h) and the command line (diskann.cpp:84-125).  The pure functions are
pinned on the CPU against closed forms and hand-worked cases; the sweep over a
GPU index is a gpu test."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, ROOT)
from distributed_diskann_amd import sweep  # noqa: E402
from distributed_diskann_amd.sweep import NNResult  # noqa: E402


def _r(recall, qps, beam=10):
    return NNResult(recall, 0, 0, 0, 0, qps, 10, beam, 1.35, 0, 0, 10, 5)


def test_statistics_closed_form():
    # average = floor(sum / n); tail = sorted[(int)(.99 * n)]
    v = np.arange(1, 201, dtype=np.uint32)[::-1]
    avg, tail = sweep.statistics(v)
    assert avg == int(sum(range(1, 201)) / 200)  # 100.5 -> 100
    assert tail == 199  # index int(.99 * 200) = 198 of 1..200
    avg, tail = sweep.statistics([7])
    assert (avg, tail) == (7, 7)
    # n = 100: the double product .99 * 100 rounds to 99.0, index 99
    avg, tail = sweep.statistics(np.arange(100, dtype=np.uint32))
    assert (avg, tail) == (49, 99)
    # n = 7: 6.93 -> index 6, the maximum
    assert sweep.statistics(np.array([5, 1, 9, 3, 3, 0, 2], np.uint32)) == (3, 9)


def test_statistics_u32_wrap():
    # the reduce is over u32 and its (int) cast goes negative; the average of
    # the negative sum over a double count, truncated, then taken as unsigned
    v = np.full(4, 0x60000000, np.uint32)  # sum 0x180000000 -> 0x80000000 -> INT_MIN
    avg, _ = sweep.statistics(v)
    assert avg == (int(-(1 << 31) / 4.0) & 0xFFFFFFFF)


def test_calculate_limits_closed_form():
    assert sweep.calculate_limits(100) == [10, 20, 30, 40, 50, 60, 70, 80, 90]
    assert sweep.calculate_limits(64) == [6, 12, 19, 25, 32, 38, 44, 51, 57]
    # duplicates removed (small bounds): 3 -> 0.3, 0.6, 0.9, 1.2, ... truncated
    assert sweep.calculate_limits(3) == [0, 1, 2]
    assert sweep.calculate_limits(0) == [0]
    # i * ub * .1 in float-then-double arithmetic: 3 * .1 is 0.30000000000000004
    # and 10 * 3 * .1 never occurs (i <= 9); a float product that rounds
    big = 16777217  # 2^24 + 1 is not a float: float(i * ub) rounds
    assert sweep.calculate_limits(big)[0] == int(16777216.0 * 0.1)


def test_parse_result_buckets():
    res = [_r(0.5, 100.0), _r(0.55, 300.0), _r(0.91, 50.0), _r(0.92, 80.0), _r(1.0, 10.0)]
    best, buckets = sweep.parse_result(res)
    by = dict(zip(buckets, best))
    # 0.5 bucket: recalls in [0.5, 0.6] -> the 300 QPS run
    assert by[0.5].QPS == 300.0
    # 0.9 bucket: [0.9, 0.93] -> 80 QPS
    assert by[0.9].QPS == 80.0
    # buckets with nothing in range are dropped: 0.6, 0.7, ... up to 0.85 have none
    assert 0.6 not in by and 0.85 not in by
    # 0.1 bucket: recall in [0.1, 0.2] -> none
    assert 0.1 not in by
    # the last bucket keeps every run with recall >= 0.99999 (no upper bound)
    assert by[0.99999].recall == 1.0
    # the buckets below it are bounded by the next bucket, so recall 1.0 misses them
    assert 0.9999 not in by and 0.99995 not in by


def test_parse_result_first_maximum():
    a, b = _r(0.95, 100.0, beam=20), _r(0.951, 100.0, beam=30)
    best, buckets = sweep.parse_result([a, b])
    assert best[buckets.index(0.95)].beamQ == 20


def test_write_to_csv_layout(tmp_path):
    g = sweep.GraphInfo("xs", "R=32", 1000, 0.5, 12.5, 32)
    res = [_r(0.95, 1000.0)]
    p = tmp_path / "r.csv"
    sweep.write_to_csv(str(p), [0.95], res, g)
    lines = p.read_text().splitlines()
    assert lines[0] == "GRAPH,Parameters,Size,Build time,Avg degree,Max degree"
    assert lines[1] == "xs,R=32,1000,0.5,12.5,32"
    assert lines[3].startswith("Num queries,Target recall,Actual recall,QPS")
    assert lines[4] == "5,0.95,0.95,1000.0,0,0,0,0,10,10,1.35"
    assert lines[5] == "" and lines[6] == ""


class _FakeIndex:
    """Answers every query with the ground truth when beam >= 20, with a
    shifted answer otherwise; counts follow the beam."""
    def __init__(self, gt, n=1000, R=32):
        self.gt, self.n, self.R, self.calls = gt, n, R, []

    def search(self, Q, k, beam, cut, limit=0, degree_limit=0, starts=(0,), trav="native", rerank=0):
        self.calls.append((k, beam, cut, limit, degree_limit))
        nq = Q.shape[0]
        ids = np.zeros((nq, k), np.uint32)
        m = min(k, self.gt.shape[1])
        ids[:, :m] = self.gt[:, :m] if beam >= 20 else self.gt[:, :m] + 100000
        c = np.full(nq, beam * 3, np.uint32)
        return ids, None, c, c // 3


def test_search_and_parse_schedule():
    nq = 8
    gt = np.arange(nq * 100, dtype=np.uint32).reshape(nq, 100)
    gd = np.tile(np.arange(100, dtype=np.float64), (nq, 1))
    fi = _FakeIndex(gt)
    rep = sweep.search_and_parse(fi, np.zeros((nq, 4), np.float32), gt, gd, k=10, log=None)
    beams = [b for b in sweep.BEAMS if b > 10]
    n_lim = len(sweep.calculate_limits(11))  # first run's average visited = beam 11
    n_deg = len(sweep.calculate_limits(32)) + 1
    assert len(rep.results) == len(beams) + n_lim * n_deg + 1
    # each run searches twice (warm-up + timed)
    assert len(fi.calls) == 2 * len(rep.results)
    assert [r.beamQ for r in rep.results[:len(beams)]] == beams
    last = rep.results[-1]
    assert (last.beamQ, last.cut, last.k) == (1000, 10.0, 10)
    assert fi.calls[-1][:3] == (100, 1000, 10.0)
    # runs with beam < 20 have recall 0, the rest 1.0
    assert {r.recall for r in rep.results} == {0.0, 1.0}
    assert rep.buckets == [0.99999]


def _write_dataset(tmp_path, n=600, nq=20, d=16, seed=3):
    from distributed_diskann_amd import formats
    rng = np.random.default_rng(seed)
    X = rng.integers(0, 256, (n, d), dtype=np.uint8)
    Q = rng.integers(0, 256, (nq, d), dtype=np.uint8)
    formats.write_vectors(str(tmp_path / "base.u8bin"), X)
    formats.write_vectors(str(tmp_path / "query.u8bin"), Q)
    D = ((Q[:, None, :].astype(np.int64) - X[None].astype(np.int64)) ** 2).sum(-1)
    ids = np.argsort(D, axis=1, kind="stable")[:, :100].astype(np.uint32)
    formats.write_truth(str(tmp_path / "gt"), ids, np.take_along_axis(D, ids.astype(np.int64), 1).astype(np.float32))
    ds = dict(dataset="tiny", nb=n, nq=nq, d=d, dtype="uint8", ds_fn="base.u8bin", qs_fn="query.u8bin", gt_fn="gt",
              basedir=str(tmp_path))
    (tmp_path / "ds.json").write_text(json.dumps(ds))
    para = {"build": True, "query": True, "save": True, "two_pass": False, "start_point_index": 0, "xita": 0.02,
            "alpha": 1.2, "L": 40, "R": 16, "Ls": 40, "max": 0, "num_neighbors_query": 100, "data_para": "ds.json"}
    (tmp_path / "para.json").write_text(json.dumps(para))
    return X, Q, ids


@pytest.mark.gpu
@pytest.mark.parametrize("original", [False, True])
def test_cli_build_and_sweep(tmp_path, original):
    from distributed_diskann_amd import formats
    X, _, _ = _write_dataset(tmp_path)
    cmd = [sys.executable, "-m", "distributed_diskann_amd", "para.json", "--sweep", "res.csv"]
    if original:
        cmd.append("--original")
    env = dict(os.environ, PYTHONPATH=ROOT + os.pathsep + os.environ.get("PYTHONPATH", ""))
    out = subprocess.run(cmd, cwd=str(tmp_path), env=env, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    g = formats.read_graph(str(tmp_path / "results" / "tiny"))
    assert g.shape == (X.shape[0], 17)
    assert (g[:, 0] <= 16).all() and (g[:, 0] > 0).all()
    lines = (tmp_path / "res.csv").read_text().splitlines()
    assert lines[0].startswith("GRAPH,")
    rows = [ln.split(",") for ln in lines[4:] if ln]
    assert rows, "no bucket rows"
    # a tiny graph is searched exactly at the largest beams
    assert float(rows[-1][2]) >= 0.99
