"""The benchmark workloads at their own shapes, against the oracle.

- one tour per launch, n in {226, 500, 1000} (random-permutation tours of a
  uniform [0,1000]^2 instance): every row of every neighborhood;
- n = 784 clustered with 256 random-insertion tours, all five neighborhoods
  in one persistent launch captured in a HIP graph and replayed: every row of
  every tour after every replay;
- 64 instances, n log-uniform in [50, 2000], batched DVND in one call (every
  execution form): final latency, rounds, moves and tour of each equal the
  oracle's DVND of that instance alone (tests/golden/dvnd_fixtures.json,
  written by tests/golden/make_dvnd_fixtures.py and re-checked by
  tests/test_oracle_dvnd_pin.py);
- a served Poisson stream of those 64 jobs plus one clustered n = 18000 job:
  every served result equals the oracle's DVND of that job alone.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

wm = pytest.importorskip("wamca_mi")

from wamca_mi import inputs as I  # noqa: E402
from wamca_mi import serve  # noqa: E402
from wamca_mi import workloads as WL  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KEY_NONE = np.iinfo(np.int64).max


@pytest.fixture(scope="module")
def eng():
    e = wm.Engine(0)
    yield e
    e.set_tile(-1)
    e.dvnd_config(4, True, False)
    e.close()


def _sha(t):
    return hashlib.sha256(np.ascontiguousarray(t, np.int32).tobytes()).hexdigest()


def _fixtures():
    with open(os.path.join(GOLDEN, "dvnd_fixtures.json")) as f:
        return json.load(f)


def _check_rows_batch(eng, slots, ref, what):
    """GPU rows of every slot and neighborhood == oracle batch rows."""
    for q, s in enumerate(slots):
        for nb, r in ref.items():
            got = eng.rows(s, nb)
            cost = r["cost"][q]
            has = cost != KEY_NONE
            bad = np.nonzero((got["cost"] != cost) | (has & ((got["i"] != r["i"][q]) |
                                                             (got["j"] != r["j"][q]))))[0]
            assert bad.size == 0, (what, q, nb, int(bad[0]), int(got["cost"][bad[0]]),
                                   int(cost[bad[0]]))


@pytest.mark.parametrize("n", [226, 500, 1000])
def test_one_tour_per_launch_sweeps(eng, n):
    xs, ys = I.uniform(n, 1, 1000)
    tours = [I.perm_tour(seed, n) for seed in range(1, 17)]
    ref, _ = O.eval_batch_xy(0x1F, 0, xs, ys, 1, tours)
    inst = eng.instance(xs, ys, 1)
    S = eng.slots(inst, len(tours))
    eng.set_tours(S, tours)
    try:
        for tile in (-1, 0, 4):
            eng.set_tile(tile)
            eng.eval_each(S, 0x1F, 0)
            eng.sync()
            _check_rows_batch(eng, S, ref, f"n={n} tile={tile}")
    finally:
        eng.set_tile(-1)
        for s in S:
            eng.destroy_slot(s)
        eng.destroy_instance(inst)


@pytest.mark.parametrize("persistent", [True, False])
def test_clustered_784_x256_graph_launch(eng, persistent):
    n, T = 784, 256
    xs, ys = I.clustered(n, 2)
    tours = [I.random_tour(xs, ys, 1, seed) for seed in range(1, T + 1)]
    ref, moves = O.eval_batch_xy(0x1F, 0, xs, ys, 1, tours)
    assert moves == T * 2444925
    inst = eng.instance(xs, ys, 1)
    S = eng.slots(inst, T)
    other = eng.slot(inst)
    try:
        eng.set_tours(S, tours)
        eng.set_tour(other, tours[0])
        tile, _ = eng.autotune(S, 0x1F, 2)
        eng.set_tile(tile)
        eng.graph_begin()
        eng.eval(S, 0x1F, 0, persistent=persistent)
        g = eng.graph_end()
        for rep in range(3):
            eng.graph_launch(g)
            eng.sync()
            _check_rows_batch(eng, S, ref, f"replay {rep} tile {tile}")
            # eager work of the program between replays
            eng.eval([other], 0x1F, 0)
            eng.merge_apply([other], 0)
            eng.sync()
        eng.graph_destroy(g)
    finally:
        eng.set_tile(-1)
        for s in S + [other]:
            eng.destroy_slot(s)
        eng.destroy_instance(inst)


def _check_dvnd(recs, tours, lat, rounds=None, what=""):
    for q, rec in enumerate(recs):
        assert int(lat[q]) == rec["latency"], (what, q, rec["n"])
        assert _sha(tours[q]) == rec["final_sha256"], (what, q, rec["n"])
        if rounds is not None:
            assert int(rounds[q]) == rec["rounds"], (what, q, rec["n"])


@pytest.mark.parametrize("chunk,graph,persistent,fused", [(4, True, False, False),
                                                          (3, True, True, False),
                                                          (1, False, False, False),
                                                          (1, True, False, True),
                                                          (4, True, False, True)])
def test_batched_64_uneven_instances(eng, chunk, graph, persistent, fused):
    recs = _fixtures()["batched_64"]
    jobs, meta = WL.dvnd_jobs(64, 3)
    insts = [eng.instance(xs, ys, 1) for xs, ys, _ in jobs]
    S = [eng.slot(i) for i in insts]
    try:
        eng.dvnd_config(chunk, graph, persistent)
        eng.dvnd_fused(fused)
        for q, (_, _, t0) in enumerate(jobs):
            assert _sha(t0) == recs[q]["start_sha256"]
        eng.set_tours(S, [j[2] for j in jobs])
        rounds = eng.dvnd(S, 0)
        _check_dvnd(recs, eng.get_tours(S), eng.latency(S), rounds, f"chunk={chunk} fused={fused}")
    finally:
        eng.dvnd_fused(False)
        eng.dvnd_config(4, True, False)
        for s in S:
            eng.destroy_slot(s)
        for i in insts:
            eng.destroy_instance(i)


def test_served_stream_with_n18000_job(eng):
    fx = _fixtures()
    if "big" not in fx:
        pytest.skip("n = 18000 record not generated")
    recs = fx["batched_64"] + [fx["big"]]
    jobs, _ = WL.dvnd_jobs(64, 3)
    big = fx["big"]
    jobs.append(WL.big_job(big["seed"], big["n"]))
    assert _sha(jobs[-1][2]) == big["start_sha256"]
    # the big job arrives among the small ones
    arr = WL.poisson_arrivals(400.0, len(jobs), 11)
    order = list(range(32)) + [64] + list(range(32, 64))
    sv = serve.run_served(eng, [jobs[q] for q in order], arr)
    lat = [0] * len(jobs)
    tours = [None] * len(jobs)
    rounds = [0] * len(jobs)
    for p, q in enumerate(order):
        lat[q], tours[q], rounds[q] = sv["latency"][p], sv["tours"][p], sv["rounds"][p]
    _check_dvnd(recs, tours, lat, rounds, "served")
