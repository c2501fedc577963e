"""The oracle (oracle/oracle.cpp) pinned against the reference's own outputs
(golden vectors written by oracle/_ref, i.e. the reference's headers) and
against what the mathematics fixes.  CPU only."""
import numpy as np
import pytest

from conftest import canon, load_golden


def test_hash64_2_golden(oracle):
    z = load_golden("hash")
    for i in range(0, 4096, 7):
        assert oracle.hash64_2(int(z["ids"][i])) == int(z["hash"][i])


def test_hash64_2_splitmix_constants(oracle):
    # the SplitMix64 finaliser, written out independently
    M = (1 << 64) - 1
    for x in [0, 1, 2, 12345, 2**31 - 1, 2**40 + 3]:
        y = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M
        y = ((y ^ (y >> 27)) * 0x94D049BB133111EB) & M
        y ^= y >> 31
        assert oracle.hash64_2(x) == y


@pytest.mark.parametrize("case", ["uint8_l2_128", "int8_l2_100", "float32_l2_20", "float32_l2_200",
                                  "float32_mips_200", "uint8_mips_128", "int8_mips_100", "float32_l2_7",
                                  "float32_mips_20"])
def test_distance_golden_bits(oracle, case):
    z = load_golden("dist")
    dt, met, d = case.split("_")
    A, B, D = z[case + "_a"], z[case + "_b"], z[case + "_d"]
    got = oracle.distances(dt, met, A, B, int(d))
    assert np.array_equal(got.view(np.uint32), D.view(np.uint32))


@pytest.mark.parametrize("case", ["uint8_l2_128", "int8_l2_100", "uint8_mips_128", "int8_mips_100"])
def test_integer_distance_closed_form(oracle, case):
    z = load_golden("dist")
    dt, met, d = case.split("_")
    A, B = z[case + "_a"].astype(np.int64), z[case + "_b"].astype(np.int64)
    exact = ((A - B) ** 2).sum(1) if met == "l2" else -(A * B).sum(1)
    got = oracle.distances(dt, met, z[case + "_a"], z[case + "_b"], int(d))
    assert np.array_equal(got, exact.astype(np.float32))


def test_f32_l2_sum_order_matters(oracle):
    # a plain left-to-right sum differs from the reference's grouped order on
    # some rows: the golden bits above therefore pin the grouping itself
    z = load_golden("dist")
    A, B, D = z["float32_l2_200_a"], z["float32_l2_200_b"], z["float32_l2_200_d"]
    seq = []
    for a, b in zip(A, B):
        s = np.float32(0)
        for x, y in zip(a, b):
            e = np.float32(x - y)
            s = np.float32(s + np.float32(e * e))
        seq.append(s)
    assert not np.array_equal(np.array(seq, np.float32).view(np.uint32), D.view(np.uint32))


def test_group_by_golden(oracle):
    z = load_golden("groupby")
    got = oracle.group_by_key(z["keys"], z["vals"])
    pos = 0
    assert len(got) == len(z["gkeys"])
    for k, ln in zip(z["gkeys"], z["glens"]):
        assert got[int(k)] == z["gvals"][pos:pos + ln].tolist()
        pos += ln


def test_random_permutation_golden(oracle):
    # parlay::random_permutation<int>(n) of the reference's own parlay, on both
    # sides of the sequential threshold and of the bucket-count steps
    z = load_golden("perm")
    ns = [int(k[1:]) for k in z.files if k.startswith("n") and k[1:].isdigit()]
    assert 8191 in ns and 8192 in ns and 300000 in ns
    for n in ns:
        p = oracle.random_permutation(n)
        assert np.array_equal(p, z[f"n{n}"]), n
        assert np.array_equal(np.sort(p), np.arange(n, dtype=np.int32))


def test_random_permutation_large_branch(oracle):
    # above 2^27 elements parlay switches its bucket-count formula; pinned by
    # the SHA-256 of the reference's permutation of 2^27 + 5 elements
    import hashlib
    z = load_golden("perm")
    p = oracle.random_permutation(int(z["big_n"]))
    assert hashlib.sha256(p.tobytes()).digest() == z["big_sha256"].tobytes()


def test_random_permutation_small_closed_form(oracle):
    # below 8192 elements: Fisher-Yates from the back, swap(A[i], A[hash64(i) % (i+1)])
    # with parlay's numerical-recipes hash64 (utilities.h:131-140), written out here
    M = (1 << 64) - 1

    def h(u):
        v = (u * 3935559000370003845 + 2691343689449507681) & M
        v ^= v >> 21
        v = (v ^ (v << 37)) & M
        v ^= v >> 4
        v = (v * 4768777513237032717) & M
        v = (v ^ (v << 20)) & M
        v ^= v >> 41
        v = (v ^ (v << 5)) & M
        return v

    for n in (2, 9, 257):
        a = list(range(n))
        for i in range(n - 1, 0, -1):
            j = h(i) % (i + 1)
            a[i], a[j] = a[j], a[i]
        assert list(oracle.random_permutation(n)) == a


def test_group_by_keeps_input_order(oracle):
    keys = [5, 3, 5, 5, 3, 9]
    vals = [10, 11, 12, 13, 14, 15]
    assert oracle.group_by_key(keys, vals) == {3: [11, 14], 5: [10, 12, 13], 9: [15]}


@pytest.mark.parametrize("Ls", [10, 20, 50, 100, 200])
def test_beam_search_matches_reference_xs(oracle, xs, Ls):
    z = xs["z"]
    P = oracle.Points(xs["X"], xs["d"])
    G = oracle.Graph(xs["G"])
    ids, dists, cmps, vis = oracle.search_all(P, G, xs["Q"], 10, Ls, cut=1.35)
    assert np.array_equal(ids, z[f"ids_{Ls}"])
    assert np.array_equal(dists.view(np.uint32), z[f"dists_{Ls}"].view(np.uint32))
    assert np.array_equal(cmps, z[f"cmps_{Ls}"])
    assert np.array_equal(vis, z[f"visited_{Ls}"])


@pytest.mark.parametrize("Ls", [10, 50, 200])
def test_beam_search_matches_reference_n0(oracle, n0, Ls):
    z = n0["z"]
    P = oracle.Points(n0["X"], n0["d"])
    G = oracle.Graph(n0["G"])
    ids, dists, cmps, vis = oracle.search_all(P, G, n0["Q"], 10, Ls, cut=1.35)
    assert np.array_equal(ids, z[f"ids_{Ls}"])
    assert np.array_equal(dists.view(np.uint32), z[f"dists_{Ls}"].view(np.uint32))
    assert np.array_equal(cmps, z[f"cmps_{Ls}"])
    assert np.array_equal(vis, z[f"visited_{Ls}"])


def test_build_mode_traces_match_reference(oracle, xs):
    z = xs["z"]
    P = oracle.Points(xs["X"], xs["d"])
    G = oracle.Graph(xs["G"])
    off = 0
    for i, pid in enumerate(z["ins_ids"]):
        (fi, fd), (vi, vd), c = oracle.beam_search(P, G, xs["X"][pid], int(pid), [0], 0, 50, 0.0, 10000, 32)
        vl = int(z["ins_vlen"][i])
        assert c == int(z["ins_cmps"][i])
        assert np.array_equal(vi, z["ins_vis_ids"][off:off + vl])
        assert np.array_equal(vd.view(np.uint32), z["ins_vis_dists"][off:off + vl].view(np.uint32))
        off += vl


def test_search_exact_on_full_beam(oracle):
    # with beam >= n on a connected graph the search returns the exact k-NN
    rng = np.random.default_rng(0)
    X = rng.standard_normal((60, 8)).astype(np.float32)
    from distributed_diskann_amd.formats import align_rows
    Xa = align_rows(X)
    P = oracle.Points(Xa, 8)
    rows = np.zeros((60, 60), np.uint32)
    rows[:, 0] = 59
    for i in range(60):
        rows[i, 1:] = [j for j in range(60) if j != i]
    G = oracle.Graph(rows)
    Q = align_rows(rng.standard_normal((5, 8)).astype(np.float32))
    ids, dists, _, _ = oracle.search_all(P, G, Q, 5, 64, cut=10.0)
    for qi in range(5):
        d = ((X.astype(np.float64) - Q[qi, :8]) ** 2).sum(1)
        d[qi] = np.inf  # the self-skip of beamSearch.h:128 drops base id == query row
        order = np.argsort(d, kind="stable")
        expect = [j for j in order if j != qi][:5] if qi != 0 else order[:5]
        # the start point 0 is never skipped (it is not a neighbor expansion)
        if qi == 0:
            expect = sorted(set(expect) | {0}, key=lambda j: (d[j] if j != 0 else ((X[0] - Q[0, :8]) ** 2).sum()))[:5]
        assert list(ids[qi]) == list(expect)


def test_empty_graph_returns_start(oracle):
    X = np.zeros((10, 16), np.float32)
    X[:, 0] = np.arange(10)
    P = oracle.Points(X, 16)
    G = oracle.Graph(np.zeros((10, 9), np.uint32))
    (fi, fd), (vi, vd), c = oracle.beam_search(P, G, X[3], -1, [0], 1, 10, 1.35, 10, 8)
    assert list(fi) == [0] and list(vi) == [0] and c == 1


# ------------------------------------------------------------- robustPrune
def _rand_points(oracle, n=300, d=16, seed=0, dtype=np.float32):
    from distributed_diskann_amd.formats import align_rows
    rng = np.random.default_rng(seed)
    if dtype == np.uint8:
        X = rng.integers(0, 256, (n, d)).astype(np.uint8)
    else:
        X = rng.standard_normal((n, d)).astype(dtype)
    return oracle.Points(align_rows(X), d), X


def test_prune_huge_alpha_keeps_R_nearest(oracle):
    P, X = _rand_points(oracle)
    G = oracle.Graph(np.zeros((P.n, 9), np.uint32))
    p = 7
    cand = np.arange(P.n, dtype=np.uint32)
    d = ((X.astype(np.float64) - X[p]) ** 2).sum(1)
    out = oracle.robust_prune(P, G, p, cand, None, alpha=1e30, add=False, R=8)
    expect = [j for j in np.argsort(d, kind="stable") if j != p][:8]
    assert list(out) == expect


def test_prune_alpha_zero_keeps_nearest_only(oracle):
    P, X = _rand_points(oracle)
    G = oracle.Graph(np.zeros((P.n, 9), np.uint32))
    p = 11
    cand = np.arange(P.n, dtype=np.uint32)
    d = ((X.astype(np.float64) - X[p]) ** 2).sum(1)
    out = oracle.robust_prune(P, G, p, cand, None, alpha=0.0, add=False, R=8)
    assert list(out) == [int([j for j in np.argsort(d, kind="stable") if j != p][0])]


def _brute_prune(X, p, cand, alpha, R):
    # index.h:84-106 restated over exact float64 distances on tiny inputs
    d = lambda a, b: float(((X[a].astype(np.float64) - X[b]) ** 2).sum())
    cs = sorted([(d(c, p), c) for c in cand])
    alive = [True] * len(cs)
    out = []
    for i, (dc, c) in enumerate(cs):
        if len(out) >= R:
            break
        if not alive[i] or c == p:
            continue
        out.append(c)
        for j in range(i + 1, len(cs)):
            if alive[j] and alpha * d(c, cs[j][1]) <= cs[j][0]:
                alive[j] = False
    return out


@pytest.mark.parametrize("alpha", [1.0, 1.2, 2.0])
def test_prune_brute_force_small(oracle, alpha):
    P, X = _rand_points(oracle, n=40, d=4, seed=3, dtype=np.uint8)
    G = oracle.Graph(np.zeros((P.n, 9), np.uint32))
    for p in range(0, 40, 5):
        cand = np.arange(40, dtype=np.uint32)
        out = oracle.robust_prune(P, G, p, cand, None, alpha=alpha, add=False, R=8,
                                  tie_mode=oracle.TIE_ID)
        assert list(out) == _brute_prune(X, p, list(range(40)), alpha, 8)


def test_prune_invariant(oracle):
    # no selected neighbor is alpha-dominated by an earlier selected one
    P, X = _rand_points(oracle, n=400, d=12, seed=5)
    G = oracle.Graph(np.zeros((P.n, 17), np.uint32))
    for p in range(0, 400, 37):
        out = oracle.robust_prune(P, G, p, np.arange(400, dtype=np.uint32), None, alpha=1.2, add=False, R=16)
        dd = lambda a, b: float(oracle.distances("float32", "l2", P.rows[[a]], P.rows[[b]], 12)[0])
        for i in range(len(out)):
            for j in range(i):
                assert not (1.2 * dd(out[j], out[i]) <= dd(p, out[i]))


# ---------------------------------------------------------- batch schedule
def _schedule_closed_form(m, max_batch):
    out, inc, count = [], 0, 0
    while count < m:
        if 2 ** inc <= max_batch:
            lo, hi = 2 ** inc - 1, min(2 ** (inc + 1), m) - 1
            count = hi
        else:
            lo, hi = count, min(count + max_batch, m)
            count += max_batch
        out.append((lo, hi))
        inc += 1
    return out


@pytest.mark.parametrize("n", [10_000, 100_000, 1_000_000, 10_000_000])
def test_schedule(oracle, n):
    s = oracle.schedule(n, 2.0, 1.0, n)
    assert s == _schedule_closed_form(n, min(n, 1_000_000))
    # the batches tile [0, n) exactly
    assert s[0] == (0, 1)
    assert all(a[1] == b[0] for a, b in zip(s, s[1:])) and s[-1][1] == n
    expected_count = {10_000: 15, 100_000: 18, 1_000_000: 21, 10_000_000: 29}[n]
    assert len(s) == expected_count


def test_schedule_bsp_xita(oracle):
    # BSP sizing over m = min(max, n) with xita (bsp_index.hpp:162-166)
    s = oracle.schedule(5000, 2.0, 0.02, 5000)
    assert s == _schedule_closed_form(5000, 100)


# ------------------------------------------------------------------- build
def test_build_random_xs_recall(oracle, xs):
    from distributed_diskann_amd.index import recall_with_ties
    P = oracle.Points(xs["X"], xs["d"])
    G = oracle.build(P, 32, 50, 1.2, tie_mode=oracle.TIE_REFERENCE)
    assert np.array_equal(canon(G.rows), xs["G"])  # deterministic restatement
    ids, _, _, _ = oracle.search_all(P, G, xs["Q"], 10, 50, cut=1.35)
    # the reference's own ground truth file (sklearn brute force)
    assert recall_with_ties(ids, xs["gt_ids"], xs["gt_d"], 10) > 0.98


def test_build_graph_invariants(oracle, n0):
    P = oracle.Points(n0["X"], n0["d"])
    for tie in (oracle.TIE_REFERENCE, oracle.TIE_ID):
        G = oracle.build(P, 32, 50, 1.2, tie_mode=tie)
        deg = G.rows[:, 0]
        assert deg.max() <= 32 and deg.min() >= 1
        for i in range(0, P.n, 97):
            nb = G.rows[i, 1:1 + deg[i]]
            assert (nb < P.n).all() and i not in nb


def test_bsp_equals_build_without_sort(oracle, n0):
    P = oracle.Points(n0["X"], n0["d"])
    a = oracle.build(P, 32, 50, 1.2, final_sort=False, tie_mode=oracle.TIE_ID)
    b = oracle.build(P, 32, 50, 1.2, final_sort=False, bsp=True, max_fraction=1.0, tie_mode=oracle.TIE_ID)
    assert np.array_equal(canon(a.rows), canon(b.rows))


def test_tie_modes_agree_on_recall(oracle, xs):
    from distributed_diskann_amd.index import recall_with_ties
    P = oracle.Points(xs["X"], xs["d"])
    r = []
    for tie in (oracle.TIE_REFERENCE, oracle.TIE_ID):
        G = oracle.build(P, 32, 50, 1.2, tie_mode=tie)
        ids, _, _, _ = oracle.search_all(P, G, xs["Q"], 10, 20, cut=1.35)
        r.append(recall_with_ties(ids, xs["gt_ids"], xs["gt_d"], 10))
    assert abs(r[0] - r[1]) < 0.005


def test_ground_truth_matches_reference_file(oracle, xs):
    P = oracle.Points(xs["X"], xs["d"])
    ids, d = oracle.ground_truth(P, xs["Q"][:200], 10)
    # the reference file stores non-squared distances
    assert np.allclose(np.sqrt(d), xs["gt_d"][:200, :10], rtol=1e-5)
    assert (ids == xs["gt_ids"][:200, :10]).mean() > 0.999


@pytest.mark.parametrize("case", ["uint8_100", "int8_100", "uint8_17", "int8_33"])
@pytest.mark.parametrize("qq", [0, 1])
def test_quantized_mips_golden_bits(oracle, case, qq):
    # the reference's own Quantized_Mips_Point::distance (mips_point.h:111-187)
    z = load_golden("qmips")
    got = oracle.quantized_mips(z[case + "_codes"], z[f"{case}_q{qq}"], float(z[case + "_max"]),
                                float(z[case + "_min"]))
    assert np.array_equal(got.view(np.uint32), z[f"{case}_out{qq}"].view(np.uint32))


def test_quantized_mips_decoding(oracle):
    # uint8 code c decodes to c * (max - min) / 255 + min, so a code row of
    # 255s against a one-hot query is max, of 0s is min; the sum is not negated
    codes = np.array([[255, 0], [0, 0]], np.uint8)
    q = np.array([[1.0, 0.0]], np.float32)
    got = oracle.quantized_mips(codes, q, 2.0, -1.0)
    assert np.allclose(got, [[2.0, -1.0]], atol=1e-6)
