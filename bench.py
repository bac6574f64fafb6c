"""Benchmark of the MI355X MLP local search.

Workload (one GPU): the full-neighborhood sweep at n = 1000 -- one synthetic
Euclidean instance (uniform integer coordinates in [0,1000]^2 from seed 1,
nint distances), 1000 random-permutation tours (seeds 1..1000, depot fixed),
each evaluated over swap, 2-opt and or-opt-1/2/3 (3,982,521 moves per tour)
with one launch per tour.  A step is one pass over the 1000 tours; inputs are
resident in HBM before the timed region.

Prints one JSON line: moves/s (value), the k_eval roofline against the LDS
operand bytes of the reference formulation, the batched (all tours in one
persistent launch) rate, and the CPU oracle timed on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "wamca-gpu-mi_amd"))

import numpy as np  # noqa: E402

# MI355X published figures used as roofline denominators
PEAK_HBM_GBS = 8000.0
CUS, LDS_BYTES_PER_CLK, CLOCK_GHZ = 256, 128, 2.4
PEAK_LDS_GBS = CUS * LDS_BYTES_PER_CLK * CLOCK_GHZ  # 78,643 GB/s

# lane-varying operand bytes per move of the reference's ADS formulation
# (swap 7 words, 2-opt 5, or-opt 5 forward / 6 backward -> 22 B average)
LDS_BYTES = {0: 28, 1: 20, 2: 22, 3: 22, 4: 22}


def moves_per_nb(N):
    return {0: (N - 4) * (N - 3) // 2, 1: ((N + 1) // 2) * (N - 3),
            2: (N - 3) * (N - 2), 3: (N - 4) * (N - 3), 4: (N - 5) * (N - 4)}


def cpu_baseline(xs, ys, tours, budget_s=12.0):
    """Oracle (single thread) per-row bests of the 5 neighborhoods on a sample."""
    from oracle import oracle as O
    W = O.weights(xs, ys, 1)
    done, moves, t0 = 0, 0, time.perf_counter()
    mp = sum(moves_per_nb(len(tours[0])).values())
    while time.perf_counter() - t0 < budget_s and done < len(tours):
        for nb in range(5):
            O.eval_rows(nb, 0, W, tours[done])
        done += 1
        moves += mp
    dt = time.perf_counter() - t0
    return {"value": moves / dt, "unit": "moves/s", "cores": 1, "kind": "port",
            "sample": f"{done} of the 1000 n=1000 tours x 5 neighborhoods "
                      f"({moves} moves, {dt:.1f} s, oracle eval_rows, 1 thread)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1000)
    ap.add_argument("--tours", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import wamca_mi
    from wamca_mi import inputs as I

    n, T = args.n, args.tours
    N = n + 1
    xs, ys = I.uniform(n, 1, 1000)
    tours = [I.perm_tour(seed, n) for seed in range(1, T + 1)]
    eng = wamca_mi.Engine(0)
    info = eng.device_info()
    inst = eng.instance(xs, ys, 1)
    slots = eng.slots(inst, T)
    eng.set_tours(slots, tours)
    eng.sync()
    mp = moves_per_nb(N)
    per_tour = sum(mp.values())

    def step():
        for s in slots:
            eng.eval([s], 0x1F, 0)

    for _ in range(args.warmup):
        step()
    eng.sync()
    eng.timing_begin(args.steps * T + 8)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    eng.sync()
    dt = time.perf_counter() - t0
    kern_ms, launches = eng.timing_end()
    value = args.steps * T * per_tour / dt
    avg_launch_s = kern_ms / 1e3 / launches

    lds_bytes = sum(mp[nb] * LDS_BYTES[nb] for nb in range(5))
    ach = lds_bytes / avg_launch_s / 1e9
    # HBM algorithmic bytes per launch: the O(N) device ADS read once
    # (xy, w int32; T, Q int32 path) plus the per-row records written
    rows = sum(wamca_mi.rows_of(nb, 0, N) for nb in range(5))
    hbm_bytes = 4 * 4 * N + 8 * rows

    # batched form: every tour in one persistent launch
    for _ in range(2):
        eng.eval(slots, 0x1F, 0, persistent=True)
    eng.sync()
    eng.timing_begin(16)
    t1 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        eng.eval(slots, 0x1F, 0, persistent=True)
    eng.sync()
    dtb = time.perf_counter() - t1
    bk_ms, bl = eng.timing_end()
    batched = reps * T * per_tour / dtb

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
                     "frac": ach / PEAK_LDS_GBS, "traffic": None,
                     "kernel": "k_eval", "avg_launch_us": avg_launch_s * 1e6,
                     "algorithmic_bytes_per_launch": lds_bytes,
                     "hbm_frac": hbm_bytes / avg_launch_s / 1e9 / PEAK_HBM_GBS},
        "batched": {"moves_per_s": batched, "tours_per_launch": T,
                    "avg_launch_ms": bk_ms / max(bl, 1),
                    "lds_frac": lds_bytes * T / (bk_ms / max(bl, 1) / 1e3) / 1e9 / PEAK_LDS_GBS},
        "device": info,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(xs, ys, tours)
    eng.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
