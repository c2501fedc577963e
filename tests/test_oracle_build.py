"""The oracle's batch_insert / build_index composition (index.h:213-341,
125-137) pinned by builds whose result the mathematics fixes.  CPU only.

Tiny cases (four 1-D uint8 points, worked by hand from index.h and
beamSearch.h; the expected rows are literals):
  * a reverse group that makes |S| + |G[t]| == R is appended (index.h:311);
  * one that makes it R + 1 is pruned with G[t] among the candidates
    (index.h:70-75, 314);
  * a reverse prune keeps an old edge of G[t] that beats the new source;
  * in the second pass of two_pass (index.h:131) a batch point is another
    batch point's target, so the forward rows must be written before the
    reverse step (index.h:298-300 before 309-322).
Closed forms on 24 random 1-D points (L >= n, so every search visits every
point of the earlier batches):
  * alpha huge, R >= n - 1: forward lists are the earlier-batch points by
    (distance, id); every reverse group is appended, so row t is its forward
    list followed by the later-batch points in insertion order;
  * alpha = 0 on L2: a forward list is the one nearest earlier-batch point;
    row t is that point followed by the later points whose nearest it is.
tests/test_oracle_mutants.py plants plausible mistakes in copies of
oracle.cpp and checks that each one fails some case here."""
import numpy as np
import pytest

from test_oracle import _schedule_closed_form


def _points(O, vals):
    X = np.zeros((len(vals), 64), np.uint8)  # d = 1, aligned to 64 bytes (point_range.h:45-50)
    X[:, 0] = vals
    return O.Points(X, 1)


def _rows(G):
    return {i: [int(v) for v in G.rows[i, 1:1 + G.rows[i, 0]]] for i in range(G.rows.shape[0])}


def _insert_all(O, vals, R, L, alpha, tie=None):
    P = _points(O, vals)
    G = O.Graph(np.zeros((len(vals), R + 1), np.uint32))
    O.batch_insert(P, G, np.arange(len(vals)), alpha, L, R=R, base=2.0, max_fraction=1.0, start=0,
                   tie_mode=O.TIE_ID if tie is None else tie)
    return _rows(G)


# ------------------------------------------------------------ tiny cases
def case_append_at_R(O):
    # batches [0,1) [1,3) [3,4); after [1,3) point 0 receives S = [1, 2]:
    # |S| + |G[0]| = 2 = R -> appended in batch order (a prune would drop 1,
    # which 2 dominates: 1.2 * d(2,1) = 1.2 <= d(0,1) = 4).  Point 3 (x=20)
    # then searches 0 -> {1, 2}, keeps 1 (it prunes 2 and 0) and 1 gets [0, 3].
    assert _insert_all(O, [10, 12, 11, 20], R=2, L=4, alpha=1.2) == {0: [1, 2], 1: [0, 3], 2: [0], 3: [1]}


def case_prune_above_R(O):
    # R = 1: S = [1, 2] for point 0 is R + 1 -> robustPrune keeps the nearest
    # (2); point 3 picks 2, whose group [3] plus G[2] = [0] is again R + 1 and
    # prunes to its nearest candidate, the old edge 0.
    assert _insert_all(O, [10, 12, 11, 20], R=1, L=4, alpha=1.2) == {0: [2], 1: [0], 2: [0], 3: [2]}


def case_reverse_prune_keeps_old_edges(O):
    # point 3 (x=7) keeps only 0; 0's group [3] plus G[0] = [1, 2] is 3 > R:
    # candidates (2, 1), (1, 4), (3, 9) -> 2 selected, 1 pruned (1.2 <= 4),
    # 3 kept (1.2 * 16 > 9).  Without G[0] in the prune the row would be [3].
    assert _insert_all(O, [10, 12, 11, 7], R=2, L=4, alpha=1.2) == {0: [2, 3], 1: [0], 2: [0], 3: [0]}


def case_two_pass_rows_before_reverse(O):
    # build_index with two_pass (alpha 1.0, then 1.2) and the final row sort.
    # Pass 1 gives {0:[1,2], 1:[0,3], 2:[0], 3:[1]}.  Pass 2: point 0 keeps
    # [2] and 2 appends 0 -> G[2] = [0, 0]; in batch {1, 2} point 1 keeps
    # [2, 3] and point 2 keeps [0]; the rows are written first (G[2] = [0]),
    # so 2's group [1] is appended: G[2] = [0, 1].  Written after the reverse
    # step, G[2] would end as [0].
    P = _points(O, [10, 12, 11, 20])
    G = O.build(P, 2, 4, 1.2, two_pass=True, tie_mode=O.TIE_ID)
    assert _rows(G) == {0: [2, 2], 1: [2, 3], 2: [0, 1], 3: [1]}


# ------------------------------------------------------------ closed forms
def _batches(n):
    return _schedule_closed_form(n, n)  # base 2, max_fraction 1


def _order(vals, t, cands):
    return sorted(cands, key=lambda j: ((int(vals[j]) - int(vals[t])) ** 2, j))


def _vals(n, seed):
    rng = np.random.default_rng(seed)
    return rng.choice(256, n, replace=False).astype(np.uint8)


def case_huge_alpha_closed_form(O):
    n = 24
    for seed in (1, 2):
        vals = _vals(n, seed)
        for R in (n - 1, n + 4):
            got = _insert_all(O, vals, R=R, L=n + 8, alpha=1e30)
            want = {}
            done = []
            batch_of = {}
            for b, (lo, hi) in enumerate(_batches(n)):
                for i in range(lo, hi):
                    batch_of[i] = b
            for lo, hi in _batches(n):
                for s in range(lo, hi):
                    want[s] = _order(vals, s, done)
                done += list(range(lo, hi))
            for t in range(n):
                want[t] = want[t] + [s for s in range(n) if batch_of[s] > batch_of[t]]
            assert got == want, (seed, R)


def case_alpha_zero_closed_form(O):
    n = 24
    for seed in (3, 4):
        vals = _vals(n, seed)
        got = _insert_all(O, vals, R=n + 4, L=n + 8, alpha=0.0)
        fwd, done = {}, []
        for lo, hi in _batches(n):
            for s in range(lo, hi):
                fwd[s] = _order(vals, s, done)[:1]
            done += list(range(lo, hi))
        want = {t: fwd[t] + [s for s in range(n) if fwd[s] == [t]] for t in range(n)}
        assert got == want, seed


CASES = [case_append_at_R, case_prune_above_R, case_reverse_prune_keeps_old_edges,
         case_two_pass_rows_before_reverse, case_huge_alpha_closed_form, case_alpha_zero_closed_form]


def test_tiny_cases_premise(oracle):
    # the hand-worked cases assume ids 0..3 sit in distinct slots of the
    # 1024-slot visited filter (beamSearch.h:66-73), so no check is lost
    slots = {oracle.hash64_2(i) & 1023 for i in range(4)}
    assert len(slots) == 4


@pytest.mark.parametrize("case", CASES, ids=[c.__name__ for c in CASES])
def test_build_closed_form(oracle, case):
    case(oracle)
