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
