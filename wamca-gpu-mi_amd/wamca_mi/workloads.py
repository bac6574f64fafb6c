"""Seeded job streams for batched and served DVND.

- ``dvnd_jobs``: instances with n = round(exp(U[ln lo, ln hi])), even indices
  uniform integer coordinates in [0,1000]^2, odd indices 20 Gaussian clusters
  (sigma 40, centres uniform in [0,2000]^2), each with its own seed, and one
  random-insertion start tour each (greedy randomized construction, alpha
  0.5, MLSolution::random of the reference, mlsolution.cpp:753-874).
  A fraction ``big_frac`` of the jobs is replaced by clustered n = big_n
  (their start tours come from the selection-based variant of the
  construction, ``inputs.random_tour_fast``).
- ``poisson_arrivals``: arrival times of a Poisson process of a given rate.
All draws come from MT19937 streams of the library's host code (mlp_host.h).
"""
from __future__ import annotations

import math

from . import inputs as I
from ._lib import lib


def _u01(seed: int, index: int) -> float:
    return (int(lib().mlp_mt_first(seed, index)) + 0.5) / 4294967296.0


def dvnd_jobs(count: int, seed: int, lo: int = 50, hi: int = 2000, big_frac: float = 0.0,
              big_n: int = 18000, rnd: int = 1):
    jobs, meta = [], []
    for q in range(count):
        big = big_frac > 0 and _u01(seed + 17, q) < big_frac
        iseed = seed * 100003 + q
        if big:
            n = big_n
            xs, ys, tour = big_job(iseed, n, rnd)
            jobs.append((xs, ys, tour))
            meta.append({"n": n, "kind": "clustered", "seed": iseed})
            continue
        n = I.loguniform_n(seed, q, lo, hi)
        if q % 2 == 0:
            xs, ys = I.uniform(n, iseed, 1000)
            kind = "uniform"
        else:
            xs, ys = I.clustered(n, iseed)
            kind = "clustered"
        # the reference's construction sorts per step (O(n^2 log n)); large
        # jobs use the selection-based variant
        tour = (I.random_tour_fast if n > 4000 else I.random_tour)(xs, ys, rnd, iseed)
        jobs.append((xs, ys, tour))
        meta.append({"n": n, "kind": kind, "seed": iseed})
    return jobs, meta


def stream_jobs_with_big(count: int, seed: int, lo: int = 50, hi: int = 2000,
                         big_frac: float = 0.01, big_n: int = 18000, rnd: int = 1):
    """``dvnd_jobs(count, seed, lo, hi, big_frac)``, with at least one job of
    n = big_n: if the draws made none, job count // 2 becomes one (instance
    seed as dvnd_jobs would give it)."""
    jobs, meta = dvnd_jobs(count, seed, lo, hi, big_frac, big_n, rnd)
    if not any(m["n"] == big_n for m in meta):
        q = count // 2
        iseed = seed * 100003 + q
        jobs[q] = big_job(iseed, big_n, rnd)
        meta[q] = {"n": big_n, "kind": "clustered", "seed": iseed}
    return jobs, meta


def big_job(iseed: int, n: int = 18000, rnd: int = 1):
    """A large stream job: clustered instance of seed iseed and the
    selection-based random-insertion start tour of the same seed."""
    xs, ys = I.clustered(n, iseed)
    return xs, ys, I.random_tour_fast(xs, ys, rnd, iseed)


def poisson_arrivals(rate: float, count: int, seed: int):
    t, out = 0.0, []
    for q in range(count):
        t += -math.log(1.0 - _u01(seed, q)) / rate
        out.append(t)
    return out
