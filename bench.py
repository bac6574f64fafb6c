"""Benchmark of the MI355X MLP local search.

Workload (one GPU): the full-neighborhood sweep at n = 1000 -- one synthetic
Euclidean instance (uniform integer coordinates in [0,1000]^2 from seed 1,
nint distances), 1000 random-permutation tours (seeds 1..1000, depot fixed),
each evaluated over swap, 2-opt and or-opt-1/2/3 (3,982,521 moves per tour)
with one launch per tour.  A step is one pass over the 1000 tours; inputs are
resident in HBM before the timed region.

Prints one JSON line: moves/s (value), the k_sweep roofline against the LDS
operand bytes of the reference formulation (average launch time from HIP
events around the timed region on the engine's stream), the batched rate (all
tours in one launch), and the CPU oracle timed on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "wamca-gpu-mi_amd"))

# MI355X published figures used as roofline denominators
PEAK_HBM_GBS = 8000.0
CUS, LDS_BYTES_PER_CLK, CLOCK_GHZ = 256, 128, 2.4
PEAK_LDS_GBS = CUS * LDS_BYTES_PER_CLK * CLOCK_GHZ  # 78,643 GB/s
PEAK_VALU_TLOPS = CUS * 4 * 16 * CLOCK_GHZ / 1e3  # 39.3 T lane-instructions/s

# lane-varying operand bytes per move of the reference's ADS formulation
# (swap 7 words, 2-opt 5, or-opt 5 forward / 6 backward -> 22 B average)
LDS_BYTES = {0: 28, 1: 20, 2: 22, 3: 22, 4: 22}
TILES = {0: "k_sweep<4,256>", 1: "k_sweep<8,512>", 2: "k_sweep<16,1024>", 3: "k_sweep<8,256>",
         4: "k_sweep<4,512>"}
# the kernels the ODC row map launches for each tile of the table
ODC_TILES = {0: "k_sweep<4,256,ODC>", 1: "k_sweep<8,512,ODC>", 3: "k_sweep<16,128,ODC>",
             4: "k_sweep<4,512,ODC>"}
# the kernels of ADS launches with swap and 2-opt (2-opt in the fixed tiles)
TWO_TILES = {1: "k_sweep<8,512,TWO>", 2: "k_sweep<12,1024,TWO>", 3: "k_sweep<8,256,TWO>"}


def ads_kernel(tile, mask):
    """The kernel an ADS launch of this tile and neighborhood mask runs."""
    return TWO_TILES[tile] if (mask & 3) == 3 and tile in TWO_TILES else TILES[tile]
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_sweep_n1000_traffic.json")


def moves_per_nb(N):
    return {0: (N - 4) * (N - 3) // 2, 1: ((N + 1) // 2) * (N - 3),
            2: (N - 3) * (N - 2), 3: (N - 4) * (N - 3), 4: (N - 5) * (N - 4)}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(xs, ys, tours, budget_s=30.0, chunk=100):
    """The oracle's per-row bests of the 5 neighborhoods (the ADS concatenation
    over the prefix form of the ADS, g++ -O3 -march=x86-64-v3), OpenMP over
    (tour x row) on the host cores the box grants this process, timed tour
    chunk by tour chunk until every tour is done or the budget is spent.
    Returns the record and the oracle's rows {nb: cost, i, j (tours, rows)}."""
    import numpy as np
    from oracle import oracle as O
    threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or 16,
                         os.cpu_count() or 1))
    rows = {nb: {"cost": [], "i": [], "j": []} for nb in range(5)}
    done = moves = 0
    t0 = time.perf_counter()
    while done < len(tours) and time.perf_counter() - t0 < budget_s:
        part, m = O.eval_batch_xy(0x1F, 0, xs, ys, 1, tours[done:done + chunk], threads)
        for nb in range(5):
            for k in ("cost", "i", "j"):
                rows[nb][k].append(part[nb][k])
        moves += m
        done += len(tours[done:done + chunk])
    dt = time.perf_counter() - t0
    rows = {nb: {k: np.concatenate(v) for k, v in r.items()} for nb, r in rows.items()}
    rec = {"value": moves / dt, "unit": "moves/s", "cores": threads, "kind": "port",
           "cpu_model": cpu_model(),
           "sample": f"{done} of the {len(tours)} n=1000 tours x 5 neighborhoods ({moves} moves, "
                     f"{dt:.1f} s): the oracle's per-row bests, g++ -O3 -march=x86-64-v3, "
                     f"OpenMP over (tour x row) on {threads} threads"}
    return rec, rows, done


def parity_check(eng, slots, rows, count):
    """GPU rows of the first `count` tours (their last evaluation) against the
    oracle's: every row's cost, and (i, j) where the row has a move."""
    import numpy as np
    none = np.iinfo(np.int64).max
    bad = 0
    for q in range(count):
        for nb in range(5):
            g = eng.rows(slots[q], nb)
            c = rows[nb]["cost"][q]
            has = c != none
            bad += int(((g["cost"] != c) | (has & ((g["i"] != rows[nb]["i"][q]) |
                                                   (g["j"] != rows[nb]["j"][q])))).sum())
    return bad


def odc_leg(eng, slots, xs, ys, tours, steps, check=64):
    """The same one-tour sweep under the ODC row map (the default of the
    bestNeighbor and MLKernel surfaces): fixed-row tiles folding each move into
    its ODC row, autotuned over the shapes of at most eight rows; the rows of
    the first `check` tours are compared with the oracle's ODC rows."""
    import numpy as np
    from oracle import oracle as O
    N = len(tours[0]) - 1
    per_tour = sum(moves_per_nb(N).values())
    tune = {}
    for t in ODC_TILES:
        eng.set_tile(t)
        eng.eval_each(slots, 0x1F, 1)
        eng.region_begin()
        eng.eval_each(slots, 0x1F, 1)
        tune[ODC_TILES[t]] = eng.region_end() * 1e3 / len(slots)
    tile = min(tune, key=tune.get)
    eng.set_tile([k for k, v in ODC_TILES.items() if v == tile][0])
    eng.eval_each(slots, 0x1F, 1)
    eng.sync()
    eng.region_begin()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.eval_each(slots, 0x1F, 1)
    region_ms = eng.region_end()
    dt = time.perf_counter() - t0
    eng.set_tile(-1)
    check = min(check, len(slots))
    ref, _ = O.eval_batch_xy(0x1F, 1, xs, ys, 1, tours[:check])
    none = np.iinfo(np.int64).max
    bad = 0
    for q in range(check):
        for nb in range(5):
            g = eng.rows(slots[q], nb)
            c = ref[nb]["cost"][q]
            has = c != none
            bad += int(((g["cost"] != c) | (has & ((g["i"] != ref[nb]["i"][q]) |
                                                   (g["j"] != ref[nb]["j"][q])))).sum())
    launches = steps * len(slots)
    avg_s = region_ms / 1e3 / launches
    return {"moves_per_s": launches * per_tour / dt, "rowmap": "ODC", "kernel": tile,
            "avg_launch_us": avg_s * 1e6, "autotune_us_per_launch": tune,
            "traffic": traffic_per_launch(tile), "valu": valu_record(tile, per_tour / avg_s),
            "parity": {"tours": check, "mismatched_rows": bad}}


def batched_n784(eng):
    """256 random-insertion tours (seeds 1..256) of one clustered n=784
    instance (20 clusters, sigma 40, centres in [0,2000]^2, seed 2), all five
    neighborhoods in one launch: the autotuner picks the tile shape, the faster
    of the persistent / one-workgroup-per-tile grids is captured in a HIP graph
    and replayed."""
    from wamca_mi import inputs as I
    n, T = 784, 256
    xs, ys = I.clustered(n, 2)
    inst = eng.instance(xs, ys, 1)
    slots = eng.slots(inst, T)
    eng.set_tours(slots, [I.random_tour(xs, ys, 1, seed) for seed in range(1, T + 1)])
    eng.sync()
    tile, tile_us = eng.autotune(slots, 0x1F, 3)
    eng.set_tile(tile)
    grid_us = {}
    for persistent in (False, True):
        eng.eval(slots, 0x1F, 0, persistent=persistent)
        eng.region_begin()
        for _ in range(3):
            eng.eval(slots, 0x1F, 0, persistent=persistent)
        grid_us[persistent] = eng.region_end() * 1e3 / 3
    persistent = min(grid_us, key=grid_us.get)
    eng.graph_begin()
    eng.eval(slots, 0x1F, 0, persistent=persistent)
    g = eng.graph_end()
    eng.graph_launch(g)
    eng.sync()
    reps = 10
    t0 = time.perf_counter()
    eng.region_begin()
    for _ in range(reps):
        eng.graph_launch(g)
    ms = eng.region_end() / reps
    dt = (time.perf_counter() - t0) / reps
    eng.graph_destroy(g)
    eng.set_tile(-1)
    per_tour = sum(moves_per_nb(n + 1).values())
    for s in slots:
        eng.destroy_slot(s)
    eng.destroy_instance(inst)
    return {"workload": "n=784 clustered, 256 random-insertion tours, 5 neighborhoods, one launch",
            "moves_per_launch": per_tour * T, "tile": ads_kernel(tile, 0x1F),
            "tile_us": {ads_kernel(k, 0x1F): v for k, v in enumerate(tile_us)},
            "persistent": persistent, "grid_us": {str(k): v for k, v in grid_us.items()},
            "graph_replay_ms": ms, "moves_per_s": per_tour * T / dt,
            "lds_frac": sum(moves_per_nb(n + 1)[nb] * LDS_BYTES[nb] for nb in range(5)) * T
            / (ms / 1e3) / 1e9 / PEAK_LDS_GBS}


def total_legs(eng):
    """TOTAL mode at n = 1001 (N = 1002, uniform [0,1000]^2 seed 5, truncated
    distances, random-insertion start tour seed 5) with the reference's
    logged block sizes at that size (576 swap / 2-opt, 896 or-opt-1): device
    evaluation of all rows * block entries, then the greedy merge over every
    entry (key transform + radix sort + merge + apply).  Median of 5."""
    from wamca_mi import inputs as I
    n = 1001
    xs, ys = I.uniform(n, 5, 1000)
    t0 = I.random_tour(xs, ys, 0, 5)
    inst = eng.instance(xs, ys, 0)
    s = eng.slot(inst)
    res = {"workload": "n=1001 uniform, one random-insertion tour, TOTAL entries + merge"}
    for nb, name, B in ((0, "swap", 576), (1, "2opt", 576), (2, "oropt1", 896)):
        ev, mg = [], []
        for rep in range(6):
            eng.set_tour(s, t0)
            eng.sync()
            eng.region_begin()
            eng.eval_total([s], nb, B)
            e_ms = eng.region_end()
            eng.region_begin()
            eng.merge_total([s])
            m_ms = eng.region_end()
            if rep:  # the first pass warms the allocator and the sort's setup
                ev.append(e_ms)
                mg.append(m_ms)
        kept = eng.applied(s, nb)
        eng.set_tour(s, t0)
        eng.eval_total([s], nb, B)
        r = eng.total(s)
        res[name] = {"block": B, "entries": len(r["cost"]),
                     "improving": int((r["cost"] < 0).sum()), "kept": len(kept),
                     "eval_ms": sorted(ev)[len(ev) // 2], "merge_ms": sorted(mg)[len(mg) // 2]}
    eng.destroy_slot(s)
    eng.destroy_instance(inst)
    return res


def dvnd_legs(eng, args):
    """DVND solves: n=100 alone; 64 uneven instances batched vs one at a time;
    a Poisson job stream served (continuous batching) vs plain."""
    import numpy as np
    from wamca_mi import inputs as I
    from wamca_mi import serve
    from wamca_mi import workloads as WL
    res = {}
    # n = 100 uniform [0,1000]^2 seed 1, nint, random-insertion start (seed 1)
    xs, ys = I.uniform(100, 1, 1000)
    t0 = I.random_tour(xs, ys, 1, 1)
    inst = eng.instance(xs, ys, 1)
    s = eng.slot(inst)
    times = []
    for _ in range(5):
        eng.set_tour(s, t0)
        eng.sync()
        t = time.perf_counter()
        rounds = eng.dvnd([s], 0)
        times.append(time.perf_counter() - t)
    res["n100_alone"] = {"ms_per_solve": min(times) * 1e3, "latency": int(eng.latency([s])[0]),
                         "rounds": int(rounds[0])}
    eng.destroy_slot(s)
    eng.destroy_instance(inst)
    fixtures = {}
    try:  # the oracle's DVND results of these workloads
        with open(os.path.join(ROOT, "tests", "golden", "dvnd_fixtures.json")) as f:
            fixtures = json.load(f)
    except (OSError, ValueError):
        pass
    # 64 instances, n log-uniform [50, 2000], half uniform / half clustered
    jobs, meta = WL.dvnd_jobs(64, 3)
    serve.run_served(eng, jobs[:8], [0.0] * 8)  # warm-up
    sv = serve.run_served(eng, jobs, [0.0] * 64, rounds_per_step=4)
    pl = serve.run_plain(eng, jobs, [0.0] * 64)
    res["batched_64"] = {
        "jobs": 64, "mean_n": float(np.mean([m["n"] for m in meta])),
        "served_solves_per_s": sv["solves_per_s"], "plain_solves_per_s": pl["solves_per_s"],
        "ratio": sv["solves_per_s"] / pl["solves_per_s"],
        "served_makespan_s": sv["makespan_s"], "plain_makespan_s": pl["makespan_s"],
        "identical_results": sv["latency"] == pl["latency"],
        "oracle_matches": (sum(int(a == r["latency"] and b == r["rounds"]) for a, b, r in
                               zip(sv["latency"], sv["rounds"], fixtures["batched_64"]))
                           if "batched_64" in fixtures else None)}
    # the same 64 jobs as one mlp_dvnd call over 64 slots, in both round forms
    insts = [eng.instance(xs, ys, 1) for xs, ys, _ in jobs]
    S = [eng.slot(i) for i in insts]
    one_call = {}
    for fused in (False, True, False, True):
        eng.dvnd_fused(fused)
        eng.set_tours(S, [j[2] for j in jobs])
        eng.sync()
        t = time.perf_counter()
        eng.dvnd(S, 0)
        dt = time.perf_counter() - t
        lat = eng.latency(S)
        key = "fused" if fused else "per_neighborhood"
        one_call[key] = {"s": min(dt, one_call.get(key, {"s": dt})["s"]),
                         "oracle_matches": (sum(int(int(a) == r["latency"]) for a, r in
                                                zip(lat, fixtures["batched_64"]))
                                            if "batched_64" in fixtures else None)}
    eng.dvnd_fused(False)
    for k in one_call:
        one_call[k]["solves_per_s"] = 64 / one_call[k]["s"]
    for x in S:
        eng.destroy_slot(x)
    for i in insts:
        eng.destroy_instance(i)
    res["batched_64"]["one_call"] = one_call
    # the same jobs with every DVND round as one persistent launch
    eng.dvnd_fused(True)
    serve.run_served(eng, jobs[:8], [0.0] * 8)  # warm-up
    sf = serve.run_served(eng, jobs, [0.0] * 64, rounds_per_step=4)
    eng.dvnd_fused(False)
    res["batched_64_fused"] = {
        "jobs": 64, "form": "one persistent launch per DVND round",
        "served_solves_per_s": sf["solves_per_s"], "plain_solves_per_s": pl["solves_per_s"],
        "ratio": sf["solves_per_s"] / pl["solves_per_s"],
        "served_makespan_s": sf["makespan_s"],
        "identical_results": sf["latency"] == pl["latency"],
        "oracle_matches": (sum(int(a == r["latency"] and b == r["rounds"]) for a, b, r in
                               zip(sf["latency"], sf["rounds"], fixtures["batched_64"]))
                           if "batched_64" in fixtures else None)}
    if args.dvnd_legs == "batched":
        return res
    # n log-uniform in [100, 1000]: all jobs at t = 0 (capacity) and a Poisson
    # stream; n in [50, 2000] with a fraction of clustered n = 18000 jobs
    for name, lo, hi, big, rate in (("burst_100_1000", 100, 1000, 0.0, 0.0),
                                    ("stream_100_1000", 100, 1000, 0.0, args.stream_rate),
                                    ("stream_50_2000_big", 50, 2000, args.stream_big_frac,
                                     args.stream_rate)):
        jobs, meta = WL.dvnd_jobs(args.stream_jobs, 4, lo=lo, hi=hi, big_frac=big)
        arr = (WL.poisson_arrivals(rate, args.stream_jobs, 40) if rate > 0
               else [0.0] * args.stream_jobs)
        sv = serve.run_served(eng, jobs, arr)
        pl = serve.run_plain(eng, jobs, arr)
        res[name] = {
            "jobs": args.stream_jobs, "rate_per_s": rate if rate > 0 else "all at t=0",
            "n": f"log-uniform [{lo}, {hi}], half uniform [0,1000]^2, half clustered; "
                 f"{sum(m['n'] == 18000 for m in meta)} jobs clustered n = 18000",
            "served": {k: sv[k] for k in ("solves_per_s", "p50_s", "p99_s", "makespan_s")},
            "plain": {k: pl[k] for k in ("solves_per_s", "p50_s", "p99_s", "makespan_s")},
            "ratio": sv["solves_per_s"] / pl["solves_per_s"],
            "identical_results": sv["latency"] == pl["latency"]}
        rec = fixtures.get("big")
        if big > 0 and rec and rec["stream"] == {"count": args.stream_jobs, "seed": 4, "lo": lo,
                                                 "hi": hi, "big_frac": big,
                                                 "job": rec["stream"]["job"]}:
            q = rec["stream"]["job"]
            res[name]["n18000_oracle_match"] = (sv["latency"][q] == rec["latency"]
                                                and sv["rounds"][q] == rec["rounds"])
    # Poisson rate sweep, n log-uniform [100, 1000] (no n = 18000 jobs): about
    # one second of arrivals per rate (60..300 jobs)
    sweep = {}
    for rate in (50.0, 200.0, 500.0, 1000.0, 2000.0, 5000.0):
        cnt = int(min(300, max(60, rate)))
        jobs, _ = WL.dvnd_jobs(cnt, 5, lo=100, hi=1000)
        arr = WL.poisson_arrivals(rate, cnt, 41)
        sv = serve.run_served(eng, jobs, arr)
        pl = serve.run_plain(eng, jobs, arr)
        sweep[str(int(rate))] = {
            "jobs": cnt,
            "served": {k: sv[k] for k in ("solves_per_s", "p50_s", "p99_s")},
            "plain": {k: pl[k] for k in ("solves_per_s", "p50_s", "p99_s")},
            "ratio": sv["solves_per_s"] / pl["solves_per_s"],
            "identical_results": sv["latency"] == pl["latency"]}
    res["rate_sweep_100_1000"] = sweep
    # the serving configuration's own stream: Poisson 50 .. 2000 jobs/s, n
    # log-uniform in [50, 2000] with 1% clustered n = 18000 jobs (at least one
    # per rate); about one second of arrivals per rate (60 .. 300 jobs)
    sweep = {}
    for rate in (50.0, 100.0, 200.0, 500.0, 1000.0, 2000.0):
        cnt = int(min(300, max(60, rate)))
        jobs, meta = WL.stream_jobs_with_big(cnt, 14, lo=50, hi=2000, big_frac=0.01)
        arr = WL.poisson_arrivals(rate, cnt, 42)
        sv = serve.run_served(eng, jobs, arr)
        pl = serve.run_plain(eng, jobs, arr)
        sweep[str(int(rate))] = {
            "jobs": cnt, "n18000_jobs": sum(m["n"] == 18000 for m in meta),
            "served": {k: sv[k] for k in ("solves_per_s", "p50_s", "p99_s", "makespan_s")},
            "plain": {k: pl[k] for k in ("solves_per_s", "p50_s", "p99_s", "makespan_s")},
            "ratio": sv["solves_per_s"] / pl["solves_per_s"],
            "identical_results": sv["latency"] == pl["latency"]}
    res["rate_sweep_50_2000_big"] = sweep
    return res


def profile_figure(kernel, key):
    """A per-launch figure of a kernel of this workload from the committed
    rocprofv3 summaries: HBM bytes per launch (FETCH_SIZE/WRITE_SIZE,
    calibrated against a copy kernel) or VALU lane-operations per move
    (SQ_INSTS_VALU x 64 / moves), if present."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        return t["kernels"][kernel][key]
    except (OSError, ValueError, KeyError):
        return None


def traffic_per_launch(kernel):
    return profile_figure(kernel, "hbm_bytes_per_launch")


def valu_record(kernel, moves_per_s):
    """VALU issue rate of the sweep (its binding resource): lane-operations
    per move from the counters x moves/s, against 256 CU x 4 SIMD x 16 lanes
    x 2.4 GHz."""
    per_move = profile_figure(kernel, "valu_lane_ops_per_move")
    if per_move is None:
        return None
    ach = per_move * moves_per_s / 1e12
    return {"lane_ops_per_move": per_move, "achieved": ach, "peak": PEAK_VALU_TLOPS,
            "unit": "T lane-instructions/s", "frac": ach / PEAK_VALU_TLOPS}


def probe_masks(eng, n, tours, tiles, masks):
    """Per-launch timing by tile shape and neighborhood mask (for counter and
    trace runs): one uniform instance (nint) with `tours` random-permutation
    tours; for every (tile, mask) one launch over all tours ("batched") and one
    over the first tour ("single"), each bracketed by HIP events on the
    engine's stream.  One JSON line per launch, in launch order, so counter or
    trace output of the same command pairs with them."""
    from wamca_mi import inputs as I
    xs, ys = I.uniform(n, 1, 1000)
    inst = eng.instance(xs, ys, 1)
    slots = eng.slots(inst, tours)
    eng.set_tours(slots, [I.perm_tour(s, n) for s in range(1, tours + 1)])
    eng.sync()
    mp = moves_per_nb(n + 1)
    for tile in tiles:
        eng.set_tile(tile)
        for mask in masks:
            for mode, sl in (("batched", slots), ("single", slots[:1])):
                eng.sync()
                eng.region_begin()
                eng.eval(sl, mask, 0)
                us = eng.region_end() * 1e3
                mv = sum(v for nb, v in mp.items() if (mask >> nb) & 1) * len(sl)
                print(json.dumps({"mode": mode, "tile": tile, "kernel": ads_kernel(tile, mask),
                                  "mask": mask,
                                  "tours": len(sl), "moves": mv, "us": round(us, 2),
                                  "gmoves_per_s": round(mv / us / 1e3, 1)}), flush=True)
    eng.set_tile(-1)


def floor_probe(eng, n=1000, reps=2000):
    """The one-tour sweep's fixed cost, split: kernels with its launch shape
    (the 4x256 grid of an n=1000 sweep, 256 threads, its kernel-argument
    table) that do nothing, then add LDS and a barrier, the dependent loads a
    tile starts with, and the row folds' atomics; then the sweep itself on one
    tour.  Back-to-back launches; for kernel traces (the durations are the
    trace's), with HIP-event averages printed beside."""
    from wamca_mi import inputs as I
    xs, ys = I.uniform(n, 1, 1000)
    inst = eng.instance(xs, ys, 1)
    s = eng.slot(inst)
    eng.set_tour(s, I.perm_tour(1, n))
    N = n + 1
    grid = ((N - 3 + 3) // 4) * ((N - 2 + 255) // 256) + ((N + 1) // 2 + 3) // 4 * ((N - 3 + 255) // 256)
    names = ("empty", "lds", "loads", "loads_atomics")
    for kind, name in enumerate(names):
        eng.launch_probe(s, kind, grid, 50)
        eng.region_begin()
        eng.launch_probe(s, kind, grid, reps)
        us = eng.region_end() * 1e3 / reps
        print(json.dumps({"probe": name, "grid": grid, "reps": reps, "us_per_launch": round(us, 3)}),
              flush=True)
    eng.set_tile(0)
    eng.eval_each([s] * 50, 0x1F, 0)
    eng.region_begin()
    eng.eval_each([s] * reps, 0x1F, 0)
    us = eng.region_end() * 1e3 / reps
    print(json.dumps({"probe": "k_sweep<4,256> n=1000", "grid": grid, "reps": reps,
                      "us_per_launch": round(us, 3)}), flush=True)
    eng.set_tile(-1)
    eng.destroy_slot(s)
    eng.destroy_instance(inst)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1000)
    ap.add_argument("--tours", type=int, default=1000)
    ap.add_argument("--tile", type=int, default=-1, help="-1: autotune over the five shapes")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-dvnd", action="store_true")
    ap.add_argument("--no-total", action="store_true")
    ap.add_argument("--dvnd-legs", choices=("all", "batched"), default="all",
                    help="batched: only the n = 100 solve and the 64-instance batch in both "
                         "forms (for trace runs)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch DVND rounds directly instead of replaying captured graphs "
                         "(for kernel-trace runs)")
    ap.add_argument("--stream-jobs", type=int, default=300)
    # served vs plain compares capacity: the stream arrives faster than either
    # path solves (plain about 250 solves/s, served about 4000); the rate
    # sweep covers the slower, arrival-limited rates
    ap.add_argument("--stream-rate", type=float, default=5000.0)
    ap.add_argument("--stream-big-frac", type=float, default=0.01,
                    help="fraction of stream jobs replaced by clustered n = 18000")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run device copies of this size (counter calibration)")
    ap.add_argument("--probe-masks", action="store_true",
                    help="only time one launch per (tile, mask) over --tours tours and over one "
                         "tour (per-launch lines for counter and trace runs)")
    ap.add_argument("--probe-tiles", default="2,4")
    ap.add_argument("--floor-probe", action="store_true",
                    help="only run the fixed-cost probe kernels and the one-tour sweep")
    ap.add_argument("--probe-mask-list", default="31,29,2,1,4")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import wamca_mi
    from wamca_mi import inputs as I

    eng = wamca_mi.Engine(0)
    if args.no_graph:
        eng.dvnd_config(4, False, False)
    if args.copy_bytes:
        eng.copy_probe(args.copy_bytes, 4)
        print(json.dumps({"copy_bytes": args.copy_bytes, "reps": 4}))
        eng.close()
        return

    if args.floor_probe:
        floor_probe(eng)
        eng.close()
        return
    if args.probe_masks:
        probe_masks(eng, args.n, args.tours, [int(t) for t in args.probe_tiles.split(",")],
                    [int(m) for m in args.probe_mask_list.split(",")])
        eng.close()
        return

    n, T = args.n, args.tours
    N = n + 1
    xs, ys = I.uniform(n, 1, 1000)
    tours = [I.perm_tour(seed, n) for seed in range(1, T + 1)]
    info = eng.device_info()
    inst = eng.instance(xs, ys, 1)
    slots = eng.slots(inst, T)
    eng.set_tours(slots, tours)
    eng.sync()
    mp = moves_per_nb(N)
    per_tour = sum(mp.values())

    # autotune (untimed): tile shape for one-tour launches
    tune = {}
    tile = args.tile
    if tile < 0:
        # every tour, as in the timed region: a shorter probe stays hotter in
        # the caches than the 1000-tour pass and can rank the shapes wrongly
        probe = slots
        for t in TILES:
            eng.set_tile(t)
            eng.eval_each(probe, 0x1F, 0)
            eng.region_begin()
            eng.eval_each(probe, 0x1F, 0)
            tune[ads_kernel(t, 0x1F)] = eng.region_end() * 1e3 / len(probe)
        tile = min(tune, key=tune.get)
        tile = [k for k in TILES if ads_kernel(k, 0x1F) == tile][0]
    eng.set_tile(tile)

    for _ in range(args.warmup):
        eng.eval_each(slots, 0x1F, 0)
    eng.sync()
    eng.region_begin()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.eval_each(slots, 0x1F, 0)
    region_ms = eng.region_end()
    dt = time.perf_counter() - t0
    eng.set_tile(-1)  # later legs use the autotuned table
    # the CPU baseline right after the timed region: its rows check the GPU
    # rows every slot still holds from its last timed launch
    cpu = parity = None
    if not args.no_cpu:
        cpu, cpu_rows, checked = cpu_baseline(xs, ys, tours)
        parity = {"tours": checked, "mismatched_rows": parity_check(eng, slots, cpu_rows, checked)}
        del cpu_rows
    launches = args.steps * T
    value = launches * per_tour / dt
    avg_launch_s = region_ms / 1e3 / launches

    lds_bytes = sum(mp[nb] * LDS_BYTES[nb] for nb in range(5))
    ach = lds_bytes / avg_launch_s / 1e9
    # compulsory HBM bytes per launch: the O(N) device ADS read (xy, w, T as
    # int32 and the 2-opt G, H) plus the per-row records written
    rows = sum(wamca_mi.rows_of(nb, 0, N) for nb in range(5))
    hbm_alg = 4 * 5 * N + 8 * rows

    out = {
        "metric": "moves evaluated/s, n=1000 full-neighborhood sweep (one solution per launch)",
        "value": value,
        "unit": "moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"n={n} uniform [0,1000]^2 nint, {T} random-permutation tours "
                               "x {swap,2opt,oropt1,oropt2,oropt3}, one tour per launch",
                   "n": n, "tours": T, "moves_per_tour": per_tour, "rowmap": "ADS",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "lds", "achieved": ach, "peak": PEAK_LDS_GBS, "unit": "GB/s",
                     "frac": ach / PEAK_LDS_GBS,
                     "traffic": traffic_per_launch(ads_kernel(tile, 0x1F)),
                     "kernel": ads_kernel(tile, 0x1F), "avg_launch_us": avg_launch_s * 1e6,
                     "algorithmic_bytes_per_launch": lds_bytes,
                     "hbm_algorithmic_bytes_per_launch": hbm_alg,
                     "hbm_frac": hbm_alg / avg_launch_s / 1e9 / PEAK_HBM_GBS,
                     "valu": valu_record(ads_kernel(tile, 0x1F), per_tour / avg_launch_s)},
        "autotune_us_per_launch": tune,
        "parity_checked": parity["tours"] if parity else 0,
        "parity": parity,
        "device": info,
    }

    out["odc_one_tour"] = odc_leg(eng, slots, xs, ys, tours, args.steps)
    if not args.no_batched:
        # batched form: every tour in one launch, largest tiles
        eng.set_tile(2)
        best = None
        for persistent in (False, True):
            eng.eval(slots, 0x1F, 0, persistent=persistent)
            eng.sync()  # the wall clock starts with the stream idle
            eng.region_begin()
            t1 = time.perf_counter()
            reps = 3
            for _ in range(reps):
                eng.eval(slots, 0x1F, 0, persistent=persistent)
            bms = eng.region_end() / reps
            dtb = (time.perf_counter() - t1) / reps
            rec = {"moves_per_s": T * per_tour / dtb, "tours_per_launch": T,
                   "kernel": "k_sweep<12,1024,TWO>", "persistent": persistent,
                   "avg_launch_ms": bms,
                   "lds_frac": lds_bytes * T / (bms / 1e3) / 1e9 / PEAK_LDS_GBS,
                   "valu": valu_record("k_sweep<12,1024,TWO>", T * per_tour / (bms / 1e3))}
            if best is None or rec["moves_per_s"] > best["moves_per_s"]:
                best = rec
        out["batched"] = best
    if not args.no_batched:
        out["batched_n784_x256"] = batched_n784(eng)
    if not args.no_dvnd:
        out["dvnd"] = dvnd_legs(eng, args)
    if not args.no_total:
        out["total_mode"] = total_legs(eng)
    if cpu is not None:
        out["cpu_baseline"] = cpu
    eng.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
