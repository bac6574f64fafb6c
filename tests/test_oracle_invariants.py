"""Oracle checks that the reference's own vectors leave open, pinned by
mathematics: three independent restatements of every move cost agree, the
row maps enumerate exactly the move sets they should, merged sets are
conflict-free with additive deltas, and DVND ends in a brute-force local
optimum."""
import itertools

import numpy as np
import pytest

from oracle import oracle as O


def _case(n, seed, rnd=1, maxc=1000):
    xs, ys = O.instance_uniform(n, seed, maxc)
    W = O.weights(xs, ys, rnd)
    return W, O.perm_tour(seed, n)


@pytest.mark.parametrize("n,seed", [(5, 1), (6, 2), (7, 3), (9, 4), (12, 5), (20, 6), (33, 7),
                                    (64, 8)])
@pytest.mark.parametrize("rowmap", [0, 1])
def test_three_cost_restatements_agree(n, seed, rowmap):
    W, t = _case(n, seed)
    for nb in range(5):
        if O.rows(nb, rowmap, n + 1) <= 0:
            continue
        m = O.enumerate_moves(nb, rowmap, W, t, ads=True, odc=True, brute=True)
        assert np.array_equal(m["ads"], m["brute"]), nb
        assert np.array_equal(m["odc"], m["brute"]), nb


def _valid_pairs(nb, N):
    k = nb - 1
    out = set()
    for i in range(1, N - 1):
        for j in range(1, N - 1):
            if nb == 0 and j >= i + 2 and j <= N - 2:
                out.add((i, j))
            elif nb == 1 and j >= i + 1 and j <= N - 2:
                out.add((i, j))
            elif nb >= 2 and i != j and i + k - 1 <= N - 2 and j + k - 1 <= N - 2:
                out.add((i, j))
    return out


@pytest.mark.parametrize("N", [6, 7, 11, 12, 40, 41])
def test_row_maps_cover_the_move_sets(N):
    W, t = _case(N - 1, N)
    for nb in range(5):
        ads = O.enumerate_moves(nb, 0, W, t, ads=False)
        pairs = list(zip(ads["i"].tolist(), ads["j"].tolist()))
        valid = _valid_pairs(nb, N)
        if nb == 1:
            # ADS 2-opt rows visit each pair, middle diagonals twice (mlk2opt.cu:168-174)
            assert set(pairs) == valid
            assert len(pairs) == ((N + 1) // 2) * (N - 3)
        else:
            assert set(pairs) == valid and len(pairs) == len(valid)
        odc = O.enumerate_moves(nb, 1, W, t, ads=False)
        op = list(zip(odc["i"].tolist(), odc["j"].tolist()))
        if nb == 0:
            assert set(op) == valid  # same set, some pairs in two rows
        elif nb >= 2:
            k = nb - 1
            missing = valid - set(op)
            # ODC or-opt never generates backward shifts by one (mlkoropt.cu:718-723)
            assert missing == {(i, j) for (i, j) in valid if i - j == 1}
            assert len(op) == (N - k - 2) ** 2


def test_ads_closed_form_matches_recurrences():
    W, t = _case(80, 3)
    T, Cm = O.ads_u32(W, t)
    N = len(t)
    T64 = T.astype(np.int64)
    S = np.cumsum(T64)
    for i in range(N):
        for j in range(N):
            if i < j:
                v = (S[j] - S[i]) - (j - i) * T64[i]
            elif i > j:
                v = (i - j) * T64[i] - (S[i - 1] - (S[j - 1] if j > 0 else 0))
            else:
                v = 0
            assert (v % 2**32) == int(Cm[i, j])
    assert int(T[-1]) == sum(int(W[t[p - 1], t[p]]) for p in range(1, N))
    assert O.latency(W, t) == int(np.sum(T64))


@pytest.mark.parametrize("rowmap", [0, 1])
def test_merge_is_conflict_free_and_additive(rowmap):
    for n, seed in [(30, 1), (100, 2), (250, 3)]:
        W, t = _case(n, seed)
        for nb in range(5):
            r = O.eval_rows(nb, rowmap, W, t)
            kept, s = O.merge(nb, r["cost"], r["i"], r["j"])
            for a, b in itertools.combinations(kept, 2):
                assert O.can_merge(nb, r["i"][a], r["j"][a], r["i"][b], r["j"][b])
            assert all(r["cost"][q] < 0 for q in kept)
            costs = [int(r["cost"][q]) for q in kept]
            assert costs == sorted(costs)
            t2 = t.copy()
            for q in kept[::-1]:  # any order: the kept moves commute
                t2 = O.apply(nb, t2, int(r["i"][q]), int(r["j"][q]))
            assert O.latency(W, t2) == O.latency(W, t) + s
            t3, c, s3 = O.step(nb, rowmap, W, t)
            assert np.array_equal(t2, t3) and s3 == s and c == len(kept)


def test_greedy_merge_definition():
    # hand-made row list: ties on cost keep the lower row first (stable sort)
    cost = np.array([-5, -9, 3, -9, -1], np.int64)
    i = np.array([1, 10, 4, 11, 30], np.int32)
    j = np.array([3, 14, 6, 20, 33], np.int32)
    kept, s = O.merge(1, cost, i, j)  # 2-opt: [10,15] vs [11,21] overlap -> row 3 dropped
    assert kept.tolist() == [1, 0, 4] and s == -15


def test_dvnd_reaches_local_optimum():
    for n, seed in [(12, 1), (25, 2), (60, 3)]:
        W, t0 = _case(n, seed)
        t, L, rounds, moves = O.dvnd(W, t0)
        assert L == O.latency(W, t) and L <= O.latency(W, t0)
        assert sorted(t[1:-1].tolist()) == list(range(1, n))
        for nb in range(5):
            m = O.enumerate_moves(nb, 0, W, t, ads=False, brute=True)
            assert (m["brute"] >= 0).all(), nb


def test_distance_function_vs_numpy_float32():
    rng = np.random.RandomState(0)
    s = np.concatenate([np.arange(0, 200000), rng.randint(0, 2**31 - 1, 20000)])
    for rnd in (0, 1):
        ref = (np.sqrt(s.astype(np.float32)) + np.float32(0.5 * rnd)).astype(np.int32)
        got = np.array([O.lib().or_dist_s(int(v), rnd) for v in s[::37]])
        assert np.array_equal(got, ref[::37])


def test_nint_differs_from_exact_rounding_above_2048():
    # s = k^2 + k has sqrt just below k + 0.5; the float path rounds it up for
    # k >= 2048, exact rounding would not -- the oracle keeps the float path.
    k = 3000
    assert O.lib().or_dist_s(k * k + k, 1) == k + 1
    assert O.lib().or_dist_s(1000 * 1000 + 1000, 1) == 1000


def _extreme_instance(n, seed, span):
    """Two far clusters; the tour alternates between them (long edges, large
    total duration) -- the worst case for the delta magnitudes."""
    rng = np.random.RandomState(seed)
    xs = np.where(np.arange(n) % 2 == 0, rng.randint(0, 20, n), span - rng.randint(0, 20, n))
    ys = rng.randint(0, span + 1, n)
    return xs.astype(np.int32), ys.astype(np.int32)


@pytest.mark.parametrize("n,seed,extreme", [(40, 1, False), (90, 2, False), (60, 3, True),
                                            (120, 4, True)])
def test_int32_exactness_bounds(n, seed, extreme):
    """The bounds the device uses to choose the int32 path (per neighborhood):
    |delta| <= 4 N maxd (swap), 3 Ttot + 7 N maxd (or-opt-k, k <= 3; the
    device adds one more N maxd of margin), N (Ttot + 6 maxd) + 3 Ttot (2-opt),
    checked against every move of random and adversarial tours."""
    span = 3000
    xs, ys = _extreme_instance(n, seed, span) if extreme else O.instance_uniform(n, seed, span)
    W = O.weights(xs, ys, 1)
    maxd = int(W.max())
    tours = [O.perm_tour(seed + q, n) for q in range(3)]
    if extreme:
        tours.append(np.array([0] + list(range(1, n)) + [0], np.int32))
    for t in tours:
        N = len(t)
        w = np.array([W[t[p - 1], t[p]] for p in range(1, N)], np.int64)
        Ttot = int(w.sum())
        for nb in range(5):
            m = O.enumerate_moves(nb, 0, W, t)
            top = int(np.abs(m["ads"]).max())
            if nb == 0:
                bound = 4 * N * maxd
            elif nb == 1:
                bound = N * (Ttot + 6 * maxd) + 3 * Ttot
            else:
                bound = 3 * Ttot + 7 * N * maxd
            assert top <= bound, (nb, top, bound)
            # the or-opt bound is not loose by orders of magnitude
            if nb >= 2 and extreme:
                assert top * 20 >= bound, (nb, top, bound)


@pytest.mark.parametrize("n,seed,extreme", [(40, 1, False), (90, 2, False), (60, 3, True),
                                            (120, 4, True)])
def test_2opt_per_move_bound(n, seed, extreme):
    """The per-move 2-opt bound the 2-opt tiles use to take the int32 path
    for a row whose moves all stay below 2^31 when the tour's bound fails:
    reversing positions i..j moves each reversed node's arrival by at most
    (T[j] - T[i]) + maxd and every later node's by at most 2 maxd, so
    |delta| <= (j - i + 1)(T[j] - T[i] + maxd) + 2 (N - 1 - j) maxd;
    checked against every move of random and adversarial tours."""
    span = 3000
    xs, ys = _extreme_instance(n, seed, span) if extreme else O.instance_uniform(n, seed, span)
    W = O.weights(xs, ys, 1)
    maxd = int(W.max())
    tours = [O.perm_tour(seed + q, n) for q in range(3)]
    if extreme:
        tours.append(np.array([0] + list(range(1, n)) + [0], np.int32))
    for t in tours:
        N = len(t)
        T = np.concatenate([[0], np.cumsum([W[t[p - 1], t[p]] for p in range(1, N)])]).astype(np.int64)
        m = O.enumerate_moves(1, 0, W, t)
        i, j = m["i"].astype(np.int64), m["j"].astype(np.int64)
        bound = (j - i + 1) * (T[j] - T[i] + maxd) + 2 * (N - 1 - j) * maxd
        assert (np.abs(m["ads"]) <= bound).all()


@pytest.mark.parametrize("n,seed", [(9, 11), (16, 12), (30, 13), (41, 14)])
@pytest.mark.parametrize("rowmap", [0, 1])
def test_oropt_k_beyond_3(n, seed, rowmap):
    """Or-opt-k for k = 4..8 (ids 5..9): the ADS and ODC closed forms equal
    brute force, the row maps cover the move set, and a DVND over ids 0..9
    ends with no improving move in any of them."""
    W, t = _case(n, seed)
    N = n + 1
    for nb in range(5, 10):
        if O.rows(nb, rowmap, N) <= 0:
            continue
        m = O.enumerate_moves(nb, rowmap, W, t, ads=True, odc=True, brute=True)
        assert np.array_equal(m["ads"], m["brute"]), nb
        assert np.array_equal(m["odc"], m["brute"]), nb
        if rowmap == 0:
            pairs = set(zip(m["i"].tolist(), m["j"].tolist()))
            assert pairs == _valid_pairs(nb, N) and len(m["i"]) == len(pairs), nb
    tf, L, rounds, _ = O.dvnd(W, O.random_tour(O.MT(seed), W, 0.5), rowmap, 0, 10)
    assert L == O.latency(W, tf)
    for nb in range(10):
        if O.rows(nb, rowmap, N) > 0:
            r = O.eval_rows(nb, rowmap, W, tf)
            assert (r["cost"] >= 0).all(), nb


@pytest.mark.parametrize("n,seed,extreme", [(14, 1, True), (20, 2, True), (40, 3, False),
                                            (90, 4, True)])
def test_delta_bound_of_64bit_keys(n, seed, extreme):
    """Every |delta| of every neighborhood id 0..14 is below N (N + 32) maxd:
    the bound instance creation checks against 2^43 (row keys delta * 2^20 +
    index) and the TOTAL merge uses as its sort bias."""
    span = 3000
    xs, ys = _extreme_instance(n, seed, span) if extreme else O.instance_uniform(n, seed, span)
    W = O.weights(xs, ys, 1)
    maxd = int(W.max())
    N = n + 1
    tours = [O.perm_tour(seed + q, n) for q in range(2)]
    if extreme:
        tours.append(np.array([0] + list(range(1, n)) + [0], np.int32))
    for t in tours:
        for nb in range(15):
            if O.rows(nb, 0, N) <= 0:
                continue
            m = O.enumerate_moves(nb, 0, W, t)
            if len(m["ads"]):
                assert int(np.abs(m["ads"]).max()) < N * (N + 32) * maxd, nb
