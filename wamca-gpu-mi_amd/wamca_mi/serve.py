"""Continuous-batching DVND serving (include/mlp_serve.h) and the job-stream
driver that compares it with the plain one-job-at-a-time GPU form.

A job is an instance (client coordinates, rounding) and a start tour; it runs
DVND to its local optimum.  ``run_stream`` replays a list of jobs with given
arrival times (seconds from the start) in real time, either through the server
(jobs arriving while others run are admitted as memory allows and share the
launches) or plainly (one job at a time, in arrival order, each alone on the
GPU), and reports per-job latency (finish - arrival), p50/p99 and solves/s.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from ._lib import JobResult, ServerConfig
from .engine import ROWMAP_ADS


class Server:
    def __init__(self, eng, hbm_budget=0, max_active=0, rowmap=ROWMAP_ADS, rounds_per_step=4,
                 max_rounds=0, big_n=0, max_bypass=0):
        """big_n: jobs of n >= big_n run in a second lane (0: 4096, < 0: one lane);
        max_bypass: later jobs that may pass a queued job that does not fit
        the free budget yet (0: 64, < 0: strict arrival order)."""
        self.eng = eng
        self._L = eng._L
        cfg = ServerConfig(int(hbm_budget), int(max_active), int(rowmap), int(rounds_per_step),
                           int(max_rounds), int(big_n), int(max_bypass))
        h = C.c_void_p()
        eng._ck(self._L.mlp_server_create(eng.h, C.byref(cfg), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self._L.mlp_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def now(self) -> float:
        return float(self._L.mlp_server_now(self.h))

    def submit(self, xs, ys, tour, rnd=1, tag=0) -> int:
        xs = np.ascontiguousarray(xs, np.int32)
        ys = np.ascontiguousarray(ys, np.int32)
        tour = np.ascontiguousarray(tour, np.int32)
        job = C.c_int64()
        self.eng._ck(self._L.mlp_server_submit(self.h, xs, ys, len(xs), rnd, tour, tag,
                                               C.byref(job)))
        return job.value

    def step(self):
        a = C.c_int32()
        q = C.c_int32()
        self.eng._ck(self._L.mlp_server_step(self.h, C.byref(a), C.byref(q)))
        return a.value, q.value

    def poll(self, cap=1024):
        buf = (JobResult * cap)()
        cnt = C.c_int32()
        self.eng._ck(self._L.mlp_server_poll(self.h, buf, cap, C.byref(cnt)))
        out = []
        for r in buf[:cnt.value]:
            tour = np.ctypeslib.as_array(r.tour, shape=(r.n + 1,)).copy()
            out.append({"job": r.job, "tag": r.tag, "latency": r.latency, "rounds": r.rounds,
                        "n": r.n, "t_submit": r.t_submit, "t_admit": r.t_admit,
                        "t_finish": r.t_finish, "tour": tour, "status": r.status})
        return out

    @staticmethod
    def job_bytes(eng, n) -> int:
        return int(eng._L.mlp_server_job_bytes(n))


def _summary(lat, span, count):
    lat = np.asarray(lat)
    return {"jobs": count, "solves_per_s": count / span if span > 0 else 0.0,
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "makespan_s": span}


def run_served(eng, jobs, arrivals, rnd=1, rowmap=ROWMAP_ADS, hbm_budget=0, max_active=0,
               rounds_per_step=4, big_n=0):
    """jobs: list of (xs, ys, tour); arrivals: seconds from the start (sorted)."""
    srv = Server(eng, hbm_budget, max_active, rowmap, rounds_per_step, big_n=big_n)
    try:
        t0 = srv.now()
        nxt, results, active, queued = 0, {}, 0, 0
        while len(results) < len(jobs):
            now = srv.now() - t0
            submitted = False
            while nxt < len(jobs) and arrivals[nxt] <= now:
                xs, ys, tour = jobs[nxt]
                srv.submit(xs, ys, tour, rnd, tag=nxt)
                nxt += 1
                submitted = True
            # idle (nothing resident, queued or just submitted): wait for the
            # next arrival; a job submitted now is stepped at once
            if not submitted and active == 0 and queued == 0 and nxt < len(jobs):
                time.sleep(min(arrivals[nxt] - now, 0.05))
                continue
            active, queued = srv.step()
            for r in srv.poll():
                results[r["tag"]] = r
        span = max(r["t_finish"] for r in results.values()) - t0
        lat = [results[q]["t_finish"] - t0 - arrivals[q] for q in range(len(jobs))]
        out = _summary(lat, span, len(jobs))
        out["latency"] = [results[q]["latency"] for q in range(len(jobs))]
        out["tours"] = [results[q]["tour"] for q in range(len(jobs))]
        out["rounds"] = [results[q]["rounds"] for q in range(len(jobs))]
        return out
    finally:
        srv.close()


def run_plain(eng, jobs, arrivals, rnd=1, rowmap=ROWMAP_ADS):
    """One job at a time in arrival order, each alone on the GPU (its own DVND
    launches), waiting for arrivals when idle."""
    t0 = time.perf_counter()
    lat, lats, tours, rounds = [], [], [], []
    for q, (xs, ys, tour) in enumerate(jobs):
        now = time.perf_counter() - t0
        if arrivals[q] > now:
            time.sleep(arrivals[q] - now)
        inst = eng.instance(xs, ys, rnd)
        s = eng.slot(inst)
        eng.set_tour(s, tour)
        r = eng.dvnd([s], rowmap)
        lats.append(int(eng.latency([s])[0]))
        tours.append(eng.get_tour(s))
        rounds.append(int(r[0]))
        eng.destroy_slot(s)
        eng.destroy_instance(inst)
        lat.append(time.perf_counter() - t0 - arrivals[q])
    span = time.perf_counter() - t0
    out = _summary(lat, span, len(jobs))
    out["latency"] = lats
    out["tours"] = tours
    out["rounds"] = rounds
    return out


def main():
    """python -m wamca_mi.serve --job KIND N SEED: one DVND job alone on the GPU
    (KIND uniform or clustered; start tour from the randomized construction,
    the selection-based variant above n = 4000).
    python -m wamca_mi.serve --batch COUNT SEED: the jobs of
    workloads.dvnd_jobs(COUNT, SEED) as one mlp_dvnd call over COUNT slots
    (--fused: each round as one persistent launch)."""
    import argparse
    import json

    from . import inputs as I
    from . import workloads as WL
    from .engine import Engine
    ap = argparse.ArgumentParser()
    g = ap.add_mutually_exclusive_group(required=True)
    g.add_argument("--job", nargs=3, metavar=("KIND", "N", "SEED"))
    g.add_argument("--batch", nargs=2, type=int, metavar=("COUNT", "SEED"))
    ap.add_argument("--no-graph", action="store_true",
                    help="launch each DVND round's kernels directly (no graph replay)")
    ap.add_argument("--fused", action="store_true")
    args = ap.parse_args()
    eng = Engine(0)
    if args.no_graph:
        eng.dvnd_config(4, False, False)
    eng.dvnd_fused(args.fused)
    if args.batch:
        count, seed = args.batch
        jobs, meta = WL.dvnd_jobs(count, seed)
        insts = [eng.instance(xs, ys, 1) for xs, ys, _ in jobs]
        slots = [eng.slot(i) for i in insts]
        eng.set_tours(slots, [j[2] for j in jobs])
        eng.sync()
        t0 = time.perf_counter()
        rounds = eng.dvnd(slots, 0)
        dt = time.perf_counter() - t0
        print(json.dumps({"batch": count, "seed": seed, "seconds": dt,
                          "max_n": max(m["n"] for m in meta), "max_rounds": int(max(rounds)),
                          "latency_sum": int(sum(int(v) for v in eng.latency(slots)))}))
        eng.close()
        return
    kind, n, seed = args.job[0], int(args.job[1]), int(args.job[2])
    xs, ys = I.clustered(n, seed) if kind == "clustered" else I.uniform(n, seed, 1000)
    tour = (I.random_tour_fast if n > 4000 else I.random_tour)(xs, ys, 1, seed)
    t0 = time.perf_counter()
    out = run_plain(eng, [(xs, ys, tour)], [0.0])
    print(json.dumps({"kind": kind, "n": n, "seed": seed, "seconds": time.perf_counter() - t0,
                      "rounds": out["rounds"][0], "latency": out["latency"][0]}))
    eng.close()


if __name__ == "__main__":
    main()
