"""Pins of the oracle's DVND loop, of its O(N)-memory prefix form and of the
committed DVND fixtures (CPU only).

The reference declares ``MLGPUTask::dvnd`` without a body
(mlgputask.h:164-167); the order its driver chains the neighborhoods is
``runWAMCA2016``'s ``for k in kMin..kMax-1: runKernel(k)``
(WamcaExperiment.hpp:414-416), each kernel evaluating, merging and applying on
the tour the previous one left (WamcaExperiment.hpp:292-365).  The oracle's
``dvnd`` repeats that pass until a pass applies nothing; here it is checked
against the same loop driven step by step from Python, and the latency it
reports against the sum of the deltas each step applied.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _sha(t):
    return hashlib.sha256(np.ascontiguousarray(t, np.int32).tobytes()).hexdigest()


def _instance(q, n):
    return O.instance_clustered(n, 70 + q) if q % 2 else O.instance_uniform(n, 70 + q)


def _driven_dvnd(W, tour, rowmap, order=(0, 1, 2, 3, 4)):
    """runWAMCA2016 order, one step per neighborhood, until a pass applies no
    move: (tour, rounds, moves, summed deltas)."""
    t = np.asarray(tour, np.int32).copy()
    rounds = moves = total = 0
    while True:
        applied = 0
        for nb in order:
            t, c, s = O.step(nb, rowmap, W, t)
            applied += c
            total += s
        rounds += 1
        moves += applied
        if applied == 0:
            return t, rounds, moves, total


@pytest.mark.parametrize("rowmap", [0, 1])
def test_dvnd_equals_step_driven_loop(rowmap):
    orders_differ = False
    for q in range(20):
        n = [30, 45, 60, 80, 101][q % 5]
        xs, ys = _instance(q, n)
        W = O.weights(xs, ys, 1)
        t0 = O.random_tour(O.MT(q + 1), W, 0.5) if q % 3 else O.perm_tour(q + 1, n)
        L0 = O.latency(W, t0)
        tour, rounds, moves, total = _driven_dvnd(W, t0, rowmap)
        ft, L, r, m = O.dvnd(W, t0, rowmap)
        assert np.array_equal(ft, tour), q
        assert (r, m) == (rounds, moves), q
        assert L == O.latency(W, tour) == L0 + total, q
        # the final tour is a local optimum of every neighborhood: no row
        # holds an improving move
        for nb in range(5):
            assert (O.eval_rows(nb, rowmap, W, ft)["cost"] >= 0).all(), (q, nb)
        # the chaining order is observable: the reverse order reaches other
        # optima, so a loop that ran the neighborhoods in another order fails
        rt = _driven_dvnd(W, t0, rowmap, order=(4, 3, 2, 1, 0))
        orders_differ = orders_differ or not np.array_equal(rt[0], ft)
    assert orders_differ


def test_dvnd_max_rounds_stops_early():
    xs, ys = O.instance_uniform(80, 5)
    W = O.weights(xs, ys, 1)
    t0 = O.perm_tour(5, 80)
    full = O.dvnd(W, t0, 0)
    assert full[2] > 3
    t = t0.copy()
    for nb in range(5):
        t, _, _ = O.step(nb, 0, W, t)
    one = O.dvnd(W, t0, 0, max_rounds=1)
    assert one[2] == 1 and np.array_equal(one[0], t)


@pytest.mark.parametrize("nb", range(7))
@pytest.mark.parametrize("rowmap", [0, 1])
def test_prefix_delta_equals_ads_and_brute(nb, rowmap):
    """The O(N) prefix form gives the ADS concatenation's value and the brute
    force L(apply) - L on every move of the row map."""
    for q, n in enumerate((12, 23, 31)):
        xs, ys = _instance(q + nb, n)
        W = O.weights(xs, ys, q % 2)
        t = O.perm_tour(3 + q + nb, n)
        mv = O.enumerate_moves(nb, rowmap, W, t, ads=True, brute=True)
        pre = np.array([O.delta_prefix(nb, xs, ys, q % 2, t, int(i), int(j))
                        for i, j in zip(mv["i"], mv["j"])], np.int64)
        assert np.array_equal(pre, mv["ads"])
        assert np.array_equal(pre, mv["brute"])


@pytest.mark.parametrize("rowmap", [0, 1])
def test_prefix_rows_and_dvnd_equal_matrix_forms(rowmap):
    for q, n in enumerate((40, 150, 301)):
        xs, ys = _instance(q, n)
        W = O.weights(xs, ys, 1)
        t = O.perm_tour(11 + q, n)
        assert O.latency_xy(xs, ys, 1, t) == O.latency(W, t)
        for nb in range(5):
            a = O.eval_rows(nb, rowmap, W, t)
            b = O.eval_rows_xy(nb, rowmap, xs, ys, 1, t, threads=3)
            for k in ("cost", "i", "j", "ctx"):
                assert np.array_equal(a[k], b[k]), (n, nb, k)
        batch, moves = O.eval_batch_xy(0x1F, rowmap, xs, ys, 1, [t, O.perm_tour(12 + q, n)])
        for nb in range(5):
            assert np.array_equal(batch[nb]["cost"][0], O.eval_rows(nb, rowmap, W, t)["cost"])
        if n <= 150:
            r1 = O.dvnd(W, t, rowmap)
            r2 = O.dvnd_xy(xs, ys, 1, t, rowmap)
            assert np.array_equal(r1[0], r2[0]) and r1[1:] == r2[1:]


def test_product_generators_equal_oracle():
    """The library's host generators (bench and serving inputs) produce the
    oracle's instances and tours."""
    from wamca_mi import inputs as I
    for n, seed in ((10, 1), (101, 2), (784, 2), (1000, 1), (2000, 77)):
        assert all(np.array_equal(a, b) for a, b in zip(I.uniform(n, seed, 1000),
                                                        O.instance_uniform(n, seed, 1000)))
        assert all(np.array_equal(a, b) for a, b in zip(I.clustered(n, seed),
                                                        O.instance_clustered(n, seed)))
        assert np.array_equal(I.perm_tour(seed, n), O.perm_tour(seed, n))
    for n, seed, rnd in ((52, 0, 0), (100, 1, 1), (300, 9, 1), (784, 5, 1)):
        xs, ys = O.instance_clustered(n, seed) if seed % 2 else O.instance_uniform(n, seed)
        W = O.weights(xs, ys, rnd)
        assert np.array_equal(I.random_tour(xs, ys, rnd, seed), O.random_tour(O.MT(seed), W, 0.5))


def _fixtures():
    with open(os.path.join(GOLDEN, "dvnd_fixtures.json")) as f:
        return json.load(f)


def test_batched_fixture_inputs_and_small_jobs():
    """Every batched job regenerates to the recorded start tour; the jobs up
    to n = 400 are recomputed by the oracle and equal their records."""
    from wamca_mi import workloads as WL
    fx = _fixtures()["batched_64"]
    jobs, meta = WL.dvnd_jobs(64, 3)
    assert len(fx) == 64
    checked = 0
    for q, ((xs, ys, t0), m, rec) in enumerate(zip(jobs, meta, fx)):
        assert (m["n"], m["kind"], m["seed"]) == (rec["n"], rec["kind"], rec["seed"]), q
        assert _sha(t0) == rec["start_sha256"], q
        assert O.latency_xy(xs, ys, 1, t0) == rec["start_latency"], q
        if m["n"] <= 400:
            ft, L, r, mv = O.dvnd_xy(xs, ys, 1, t0)
            assert (L, r, mv, _sha(ft)) == (rec["latency"], rec["rounds"], rec["moves"],
                                           rec["final_sha256"]), q
            checked += 1
    assert checked >= 30


def test_big_fixture_inputs():
    fx = _fixtures()
    if "big" not in fx:
        pytest.skip("n = 18000 record not generated")
    from wamca_mi import workloads as WL
    rec = fx["big"]
    xs, ys, t0 = WL.big_job(rec["seed"], rec["n"])
    assert _sha(t0) == rec["start_sha256"]
    assert O.latency_xy(xs, ys, 1, t0) == rec["start_latency"]
    assert rec["latency"] < rec["start_latency"] and rec["rounds"] > 100
