"""Writes tests/golden/dvnd_fixtures.json: the oracle's DVND results for the
batched and served workloads, whose full-size runs are too long to recompute
inside a GPU test.

- "batched_64": the 64 jobs of ``workloads.dvnd_jobs(64, 3)`` (n log-uniform
  in [50, 2000], half uniform [0,1000]^2, half clustered, one random-insertion
  start tour each, nint distances);
- "big": the first n = 18000 clustered job of the served stream
  ``workloads.dvnd_jobs(300, 4, lo=50, hi=2000, big_frac=0.01)`` (job 232,
  instance seed 400244, selection-based start tour).

Each record holds the job's n, kind and seed, sha256 of the start tour (a
check that the inputs are regenerated unchanged), and the oracle's DVND over
swap, 2-opt, or-opt-1/2/3 with the ADS row map to a local optimum: final
latency, rounds, applied moves and sha256 of the final tour.  The oracle runs
in its O(N)-memory prefix form (``oracle.dvnd_xy``), OpenMP over rows.

Usage: python tests/golden/make_dvnd_fixtures.py [--part batched|big|all]
       [--threads T]   (the big job takes about an hour on 8 host cores)
"""
import argparse
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "wamca-gpu-mi_amd"))

import numpy as np  # noqa: E402

from oracle import oracle as O  # noqa: E402
from wamca_mi import workloads as WL  # noqa: E402

OUT = os.path.join(HERE, "dvnd_fixtures.json")
BIG_STREAM = (300, 4, 50, 2000, 0.01)  # count, seed, lo, hi, big_frac
BIG_JOB = 232


def sha(t) -> str:
    return hashlib.sha256(np.ascontiguousarray(t, np.int32).tobytes()).hexdigest()


def record(xs, ys, tour, meta, threads):
    t0 = time.time()
    ft, L, rounds, moves = O.dvnd_xy(xs, ys, 1, tour, 0, 0, 5, threads)
    return {"n": int(meta["n"]), "kind": meta["kind"], "seed": int(meta["seed"]),
            "start_sha256": sha(tour), "start_latency": O.latency_xy(xs, ys, 1, tour),
            "latency": L, "rounds": rounds, "moves": moves, "final_sha256": sha(ft),
            "oracle_s": round(time.time() - t0, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--part", default="all", choices=("batched", "big", "all"))
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args()
    out = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            out = json.load(f)
    out["_doc"] = ("oracle DVND (ADS row map, swap/2-opt/or-opt-1..3, nint) of the batched "
                   "64-job workload and of the served stream's first n=18000 job; written by "
                   "make_dvnd_fixtures.py")
    if a.part in ("batched", "all"):
        jobs, meta = WL.dvnd_jobs(64, 3)
        out["batched_64"] = [record(*jobs[q], meta[q], a.threads) for q in range(64)]
        with open(OUT, "w") as f:
            json.dump(out, f, indent=1)
    if a.part in ("big", "all"):
        count, seed, lo, hi, frac = BIG_STREAM
        # the big job is regenerated alone, exactly as dvnd_jobs builds job 232
        assert WL._u01(seed + 17, BIG_JOB) < frac
        iseed = seed * 100003 + BIG_JOB
        xs, ys, tour = WL.big_job(iseed)
        rec = record(xs, ys, tour, {"n": len(xs), "kind": "clustered", "seed": iseed}, a.threads)
        rec["stream"] = {"count": count, "seed": seed, "lo": lo, "hi": hi, "big_frac": frac,
                         "job": BIG_JOB}
        out["big"] = rec
        with open(OUT, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps({k: (len(v) if isinstance(v, list) else v) for k, v in out.items()})[:2000])


if __name__ == "__main__":
    main()
