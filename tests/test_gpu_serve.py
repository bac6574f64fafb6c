"""The continuous-batching server (include/mlp_serve.h): every job's result
is the oracle's DVND result for that job alone, whatever ran beside it, under
arrival patterns that make jobs join running batches, under a memory budget
that holds jobs back, and with an active-job limit."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

wm = pytest.importorskip("wamca_mi")
from wamca_mi import serve  # noqa: E402


@pytest.fixture(scope="module")
def eng():
    e = wm.Engine(0)
    yield e
    e.close()


def _jobs(ns, seed0):
    jobs, refs = [], []
    for q, n in enumerate(ns):
        if q % 2:
            xs, ys = O.instance_clustered(n, seed0 + q)
        else:
            xs, ys = O.instance_uniform(n, seed0 + q)
        W = O.weights(xs, ys, 1)
        t0 = O.random_tour(O.MT(seed0 + q), W, 0.5)
        jobs.append((xs, ys, t0))
        refs.append(O.dvnd(W, t0, 0))
    return jobs, refs


def _check(out, refs):
    for q, (tref, Lref, rref, _) in enumerate(refs):
        assert out["latency"][q] == Lref, q
        assert np.array_equal(out["tours"][q], tref), q
        assert out["rounds"][q] == rref, q


@pytest.mark.parametrize("max_active,budget_jobs", [(0, 0), (3, 0), (0, 2)])
def test_served_equals_alone(eng, max_active, budget_jobs):
    ns = [60, 200, 90, 350, 140, 75, 260, 120, 500, 55]
    jobs, refs = _jobs(ns, 7000 + max_active + budget_jobs)
    arrivals = [0.0, 0.0, 0.001, 0.002, 0.002, 0.004, 0.006, 0.006, 0.008, 0.01]
    budget = 0
    if budget_jobs:
        budget = budget_jobs * max(serve.Server.job_bytes(eng, n) for n in ns)
    out = serve.run_served(eng, jobs, arrivals, hbm_budget=budget, max_active=max_active)
    _check(out, refs)
    assert out["p99_s"] >= out["p50_s"] >= 0


def test_plain_equals_oracle(eng):
    jobs, refs = _jobs([80, 150, 230], 9100)
    out = serve.run_plain(eng, jobs, [0.0, 0.0, 0.0])
    _check(out, refs)


def test_server_rejects_oversized_job(eng):
    srv = serve.Server(eng, hbm_budget=4096)
    try:
        xs, ys = O.instance_uniform(100, 1)
        t = O.perm_tour(1, 100)
        with pytest.raises(wm.MLPError):
            srv.submit(xs, ys, t)
    finally:
        srv.close()


def test_two_lanes_equal_alone(eng):
    """Jobs at or above big_n run in the second lane (own context and stream);
    every result still equals the oracle's DVND of the job alone."""
    ns = [60, 700, 90, 1500, 140, 75, 900, 120]
    jobs, refs = _jobs(ns, 8300)
    arrivals = [0.0, 0.0, 0.001, 0.001, 0.002, 0.003, 0.003, 0.004]
    out = serve.run_served(eng, jobs, arrivals, big_n=600)
    _check(out, refs)


def _drain(srv, count):
    res = {}
    while len(res) < count:
        srv.step()
        for r in srv.poll():
            res[r["tag"]] = r
    return res


def test_invalid_jobs_refused_valid_ones_finish(eng):
    """A job that instance or tour set-up would refuse is refused at submit
    (nothing queued), so it cannot block or leak the jobs around it."""
    jobs, refs = _jobs([70, 130, 90], 9300)
    srv = serve.Server(eng)
    try:
        srv.submit(*jobs[0], tag=0)
        xs, ys, t = jobs[1]
        bad_t = t.copy()
        bad_t[3] = bad_t[4]  # not a permutation
        with pytest.raises(wm.MLPError) as e:
            srv.submit(xs, ys, bad_t, tag=10)
        assert e.value.status == -1
        wide = xs.copy()
        wide[5] = int(xs.min()) + 40000  # coordinate span above 32767
        with pytest.raises(wm.MLPError) as e:
            srv.submit(wide, ys, t, tag=11)
        assert e.value.status == -4
        # N (N + 32) maxd >= 2^43: 14000 clients spread over the full range
        bx, by = O.instance_uniform(14000, 3, 32767)
        with pytest.raises(wm.MLPError) as e:
            srv.submit(bx, by, O.perm_tour(3, 14000), tag=12)
        assert e.value.status == -4
        srv.submit(*jobs[1], tag=1)
        srv.submit(*jobs[2], tag=2)
        res = _drain(srv, 3)
        assert sorted(res) == [0, 1, 2]
        for q, (tref, Lref, rref, _) in enumerate(refs):
            assert res[q]["status"] == 0
            assert res[q]["latency"] == Lref and res[q]["rounds"] == rref, q
            assert np.array_equal(res[q]["tour"], tref), q
    finally:
        srv.close()


def test_negative_coordinates_served(eng):
    """Instances with negative coordinates (the reference's gil262.tsp) are
    translated on the device; distances, and so every result, are unchanged."""
    from helpers import load_tsplib_coords
    xs, ys = load_tsplib_coords("09_gil262.tsp")
    assert xs.min() < 0 and ys.min() < 0
    W = O.weights(xs, ys, 1)
    t0 = O.random_tour(O.MT(3), W, 0.5)
    ref = O.dvnd(W, t0, 0)
    out = serve.run_served(eng, [(xs, ys, t0)], [0.0])
    assert out["latency"][0] == ref[1] and out["rounds"][0] == ref[2]
    assert np.array_equal(out["tours"][0], ref[0])


def test_blocked_job_does_not_hold_small_jobs(eng):
    """The oldest queued job does not fit the free budget while a large job
    runs; smaller jobs behind it are admitted past it (up to max_bypass), and
    every result equals the oracle's."""
    ns = [1200, 1100, 60, 70, 80, 90]
    jobs, refs = _jobs(ns, 9400)
    budget = serve.Server.job_bytes(eng, 1200) + 4 * serve.Server.job_bytes(eng, 100)
    srv = serve.Server(eng, hbm_budget=budget, big_n=-1)
    try:
        for q, job in enumerate(jobs):
            srv.submit(*job, tag=q)
        srv.step()  # admits job 0 and the small jobs, not job 1
        res = {}
        for r in srv.poll():
            res[r["tag"]] = r
        res.update(_drain(srv, len(jobs) - len(res)))
        order = sorted(res, key=lambda q: res[q]["t_admit"])
        assert order.index(1) > order.index(2)  # a small job went first
        for q, (tref, Lref, rref, _) in enumerate(refs):
            assert res[q]["latency"] == Lref and np.array_equal(res[q]["tour"], tref), q
    finally:
        srv.close()


def test_served_fused_rounds_equal_alone(eng):
    """Served jobs whose rounds run as one persistent launch each: every
    result equals the oracle's DVND of the job alone."""
    ns = [60, 200, 90, 350, 140, 75, 260, 120, 500, 55]
    jobs, refs = _jobs(ns, 9500)
    arrivals = [0.0, 0.0, 0.001, 0.002, 0.002, 0.004, 0.006, 0.006, 0.008, 0.01]
    eng.dvnd_fused(True)
    try:
        out = serve.run_served(eng, jobs, arrivals)
    finally:
        eng.dvnd_fused(False)
    _check(out, refs)
