"""Batched DVND on the GPU against the oracle's DVND (oracle/mlp_oracle.cpp
or_dvnd: rounds of swap, 2-opt, or-opt-1/2/3, each evaluated, merged and
applied in turn, until a round applies nothing).

Covers every execution form of the rounds (chunk size, HIP-graph replay,
persistent grids), both row maps, max_rounds, the stepwise API a serving loop
uses (slots joining a running batch between chunks), and instances of uneven
n in one batch.  Final tours, latencies and round counts must be identical to
the oracle's for every slot, alone or in a batch.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

wm = pytest.importorskip("wamca_mi")


@pytest.fixture(scope="module")
def eng():
    e = wm.Engine(0)
    yield e
    e.dvnd_config(4, True, False)
    e.close()


def _jobs(ns, seed0, rnd=1):
    jobs = []
    for q, n in enumerate(ns):
        if q % 2:
            xs, ys = O.instance_clustered(n, seed0 + q)
        else:
            xs, ys = O.instance_uniform(n, seed0 + q)
        W = O.weights(xs, ys, rnd)
        t0 = O.random_tour(O.MT(seed0 + q), W, 0.5)
        jobs.append((xs, ys, W, t0))
    return jobs


def _run(eng, jobs, rowmap, max_rounds=0, rnd=1):
    insts = [eng.instance(xs, ys, rnd) for xs, ys, _, _ in jobs]
    S = [eng.slot(i) for i in insts]
    eng.set_tours(S, [j[3] for j in jobs])
    rounds = eng.dvnd(S, rowmap, max_rounds)
    out = [(eng.get_tour(s), int(eng.latency([s])[0]), int(rounds[q])) for q, s in enumerate(S)]
    for s in S:
        eng.destroy_slot(s)
    for i in insts:
        eng.destroy_instance(i)
    return out


@pytest.mark.parametrize("chunk,graph,persistent", [(1, False, False), (4, True, False),
                                                    (3, True, True), (8, False, True)])
@pytest.mark.parametrize("rowmap", [0, 1])
@pytest.mark.parametrize("groups", [1, 2])
def test_dvnd_forms_match_oracle(eng, chunk, graph, persistent, rowmap, groups):
    # groups: one stream, or the slots split over two streams (the persistent
    # form keeps one)
    jobs = _jobs([40, 97, 180, 333, 611], 900 + rowmap)
    eng.dvnd_config(chunk, graph, persistent)
    eng.dvnd_groups(groups)
    try:
        got = _run(eng, jobs, rowmap)
    finally:
        eng.dvnd_config(4, True, False)
        eng.dvnd_groups(2)
    for q, (xs, ys, W, t0) in enumerate(jobs):
        tref, Lref, rref, _ = O.dvnd(W, t0, rowmap)
        assert got[q][1] == Lref, (q, chunk, graph, persistent)
        assert np.array_equal(got[q][0], tref)
        assert got[q][2] == rref


@pytest.mark.parametrize("rowmap", [0, 1])
def test_dvnd_groups_uneven_and_converged(eng, rowmap):
    """Two slot groups that advance independently (each group's convergence
    check follows its own chunk of rounds): one large slot in the first group,
    small ones in the second, one of which starts at its local optimum (one
    round that applies nothing retires it at its group's first check)."""
    jobs = _jobs([700, 12, 35, 60], 1300 + rowmap)
    xs, ys, W, t0 = jobs[3]
    jobs[3] = (xs, ys, W, O.dvnd(W, t0, rowmap)[0])
    eng.dvnd_groups(2)
    got = _run(eng, jobs, rowmap)
    for q, (xs, ys, W, t0) in enumerate(jobs):
        tref, Lref, rref, _ = O.dvnd(W, t0, rowmap)
        assert (got[q][1], got[q][2]) == (Lref, rref), q
        assert np.array_equal(got[q][0], tref), q
    assert got[3][2] == 1 and np.array_equal(got[3][0], jobs[3][3])


@pytest.mark.parametrize("max_rounds", [1, 2, 5])
def test_dvnd_max_rounds(eng, max_rounds):
    jobs = _jobs([120, 260], 77)
    got = _run(eng, jobs, 0, max_rounds)
    for q, (xs, ys, W, t0) in enumerate(jobs):
        tref, Lref, rref, _ = O.dvnd(W, t0, 0, max_rounds)
        assert got[q][1] == Lref and np.array_equal(got[q][0], tref) and got[q][2] == rref


def test_dvnd_stepwise_join_equals_alone(eng):
    """Slots that join a running batch between chunks (as a serving loop
    admits jobs) end exactly where they end alone."""
    jobs = _jobs([150, 64, 300, 90, 500, 210], 4242)
    insts = [eng.instance(xs, ys, 1) for xs, ys, _, _ in jobs]
    S = [eng.slot(i) for i in insts]
    try:
        running = []
        pending = list(range(len(jobs)))
        finished = {}
        step = 0
        while pending or running:
            # admit up to two new jobs every other step
            if pending and step % 2 == 0:
                new = pending[:2]
                pending = pending[2:]
                eng.set_tours([S[q] for q in new], [jobs[q][3] for q in new])
                eng.dvnd_start([S[q] for q in new])
                running += new
            eng.dvnd_rounds([S[q] for q in running], 3, 0)
            act, rnd = eng.dvnd_status([S[q] for q in running])
            still = []
            for k, q in enumerate(running):
                if act[k]:
                    still.append(q)
                else:
                    finished[q] = (eng.get_tour(S[q]), int(eng.latency([S[q]])[0]), int(rnd[k]))
            running = still
            step += 1
        for q, (xs, ys, W, t0) in enumerate(jobs):
            tref, Lref, rref, _ = O.dvnd(W, t0, 0)
            assert finished[q][1] == Lref and np.array_equal(finished[q][0], tref)
            assert finished[q][2] == rref
    finally:
        for s in S:
            eng.destroy_slot(s)
        for i in insts:
            eng.destroy_instance(i)


def test_dvnd_many_slots_list_on_device(eng):
    """More slots than the inline kernel-argument lists hold (device lists,
    captured in the graph after their first use)."""
    xs, ys = O.instance_uniform(60, 5)
    W = O.weights(xs, ys, 1)
    inst = eng.instance(xs, ys, 1)
    tours = [O.random_tour(O.MT(1000 + q), W, 0.5) for q in range(150)]
    S = [eng.slot(inst) for _ in tours]
    try:
        eng.set_tours(S, tours)
        rounds = eng.dvnd(S, 0)
        for q in range(0, 150, 7):
            tref, Lref, rref, _ = O.dvnd(W, tours[q], 0)
            assert int(eng.latency([S[q]])[0]) == Lref
            assert np.array_equal(eng.get_tour(S[q]), tref)
            assert int(rounds[q]) == rref
    finally:
        for s in S:
            eng.destroy_slot(s)
        eng.destroy_instance(inst)


def test_standalone_eval_after_dvnd(eng):
    """Row buffers written inside DVND rounds (parity committed on the
    device) are followed correctly by standalone evals and merges."""
    xs, ys = O.instance_uniform(150, 31)
    W = O.weights(xs, ys, 1)
    inst = eng.instance(xs, ys, 1)
    s = eng.slot(inst)
    try:
        for rowmap_dvnd, rowmap_eval in ((0, 0), (1, 0), (0, 1), (1, 1)):
            t0 = O.random_tour(O.MT(31 + rowmap_dvnd), W, 0.5)
            eng.set_tour(s, t0)
            eng.dvnd([s], rowmap_dvnd, 3)
            t1 = O.perm_tour(5 + rowmap_eval, 150)
            eng.set_tour(s, t1)
            for rep in range(2):
                eng.eval([s], 0x1F, rowmap_eval)
                for nb in range(5):
                    got = eng.rows(s, nb)
                    ref = O.eval_rows(nb, rowmap_eval, W, t1)
                    assert np.array_equal(got["cost"], ref["cost"]), (rowmap_dvnd, rowmap_eval, nb)
            eng.merge_apply([s], 2)
            t2, _, _ = O.step(2, rowmap_eval, W, t1)
            assert np.array_equal(eng.get_tour(s), t2)
    finally:
        eng.destroy_slot(s)
        eng.destroy_instance(inst)


def test_cached_round_graph_follows_neighborhood_count(eng):
    """A DVND chunk captured as a graph with one neighborhood set is not
    replayed after the set changes: slots with buffers for 15 ids run DVND
    over 10 neighborhoods until the chunk graph is in use, then over 5."""
    eng.set_neighborhoods(15)
    try:
        xs, ys = O.instance_uniform(120, 44)
        W = O.weights(xs, ys, 1)
        tours = [O.random_tour(O.MT(440 + q), W, 0.5) for q in range(3)]
        inst = eng.instance(xs, ys, 1)
        S = [eng.slot(inst) for _ in tours]
        eng.dvnd_config(1, True, False)  # one round per chunk: the graph is reused at once
        for nbs in (10, 5):
            eng.dvnd_neighborhoods(nbs)
            eng.set_tours(S, tours)
            rounds = eng.dvnd(S, 0)
            for q, s in enumerate(S):
                tref, Lref, rref, _ = O.dvnd(W, tours[q], 0, nbs=nbs)
                assert int(eng.latency([s])[0]) == Lref, (nbs, q)
                assert int(rounds[q]) == rref, (nbs, q)
                assert np.array_equal(eng.get_tour(s), tref), (nbs, q)
        for s in S:
            eng.destroy_slot(s)
        eng.destroy_instance(inst)
    finally:
        eng.dvnd_neighborhoods(5)
        eng.set_neighborhoods(5)
        eng.dvnd_config(4, True, False)


@pytest.mark.parametrize("chunk,graph", [(1, False), (4, True), (3, True)])
def test_fused_rounds_match_oracle(eng, chunk, graph):
    """Each DVND round as one persistent launch (a queue of every slot's eval
    tiles and merges with per-slot dependencies): final tours, latencies and
    round counts equal the oracle's for every slot of an uneven batch."""
    jobs = _jobs([8, 40, 97, 180, 333, 611, 1200], 950 + chunk)
    eng.dvnd_config(chunk, graph, False)
    eng.dvnd_fused(True)
    try:
        out = _run(eng, jobs, 0)
    finally:
        eng.dvnd_fused(False)
        eng.dvnd_config(4, True, False)
    for (xs, ys, W, t0), (tour, L, rounds) in zip(jobs, out):
        tref, Lref, rref, _ = O.dvnd(W, t0, 0)
        assert L == Lref and rounds == rref
        assert np.array_equal(tour, tref)


def test_fused_rounds_after_staging_growth(eng):
    """Fused rounds on a few slots, then a larger batch whose tour upload grows
    the context's staging buffers, then fused rounds again: every slot's
    result equals the oracle's (the fused rounds' counters outlive the
    upload buffers' growth)."""
    eng.dvnd_fused(True)
    try:
        for ns, seed in (([50, 90, 130], 990), ([40 + 23 * q for q in range(24)], 991)):
            jobs = _jobs(ns, seed)
            out = _run(eng, jobs, 0)
            for (xs, ys, W, t0), (tour, L, rounds) in zip(jobs, out):
                tref, Lref, rref, _ = O.dvnd(W, t0, 0)
                assert L == Lref and rounds == rref
                assert np.array_equal(tour, tref)
    finally:
        eng.dvnd_fused(False)
