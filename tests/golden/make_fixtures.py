"""Extract golden vectors from the reference's published experiment logs.

Reads ``/root/reference/results/*.txt`` (text logs written by the reference's
WAMCA 2016 driver) and writes ``tests/golden/results_logs.json``: for every
logged start tour, its client sequence and uint32 latency, and for each of the
SWAP / 2OPT / OROPT1 kernels the values the log prints:

* ``best``     -- "after sort 0 (i,j) cost": first record after the stable
                  sort of the per-row list (the global best move);
* ``graph``    -- "Graph V/M": improving records / records;
* ``partial``  -- "partial GPU-GPU improvement=S count=C": merged set;
* ``total``    -- the TOTAL (per-thread-slot) section: "kernel moveElems=E",
                  "after sort 0", "Graph", and both the GPU-GPU and GPU-CPU
                  merged (S, C) pairs.

Run it in a container where the reference is mounted; the JSON is committed.
"""
import json
import os
import re

REF = "/root/reference/results"
INST = {
    "results-inst100.txt": "02_kroD100.tsp",
    "results-inst226.txt": "03_pr226.tsp",
    "results-inst300.txt": "04_lin318.tsp",
    "results-inst500.txt": "05_TRP-S500-R1.tsp",
    "results-inst600.txt": "06_d657.tsp",
    "results-inst700.txt": "07_rat784.tsp",
    "results-inst1000.txt": "08_TRP-S1000-R1.tsp",
    "results-inst1000-bug-or1.txt": "08_TRP-S1000-R1.tsp",
}
NB = {"SWAP": 0, "2OPT": 1, "OROPT1": 2, "OROPT2": 3, "OROPT3": 4}

re_seed = re.compile(r"RAND_SEED\s*:\s*(\d+)")
re_sol = re.compile(r"\* Solution #(\d+)")
re_tour = re.compile(r"^<([0-9,]+)> = (\d+) \(TOUR\)")
re_init = re.compile(r"initializing kernel (\d+) with &t?kernel:\S+ (\w+) TOTAL=(\d)")
re_elems = re.compile(r"kernel moveElems=(\d+)!")
re_blck = re.compile(r"grid\(1,(\d+),1\)\s+blck\((\d+),1,1\)")
re_best = re.compile(r"after sort 0 \((\d+),(\d+)\) (-?\d+)")
re_graph = re.compile(r"Graph\s+(\d+)/(\d+)")
re_imp = re.compile(r"(partial|total) GPU-(GPU|CPU) improvement=(-?\d+) count=(\d+)")


def parse(path):
    sols = []
    seed = None
    cur = None
    kern = None
    for line in open(path):
        line = line.rstrip("\n")
        m = re_seed.search(line)
        if m:
            seed = int(m.group(1))
            continue
        m = re_sol.search(line)
        if m:
            cur = {"index": int(m.group(1)), "kernels": {}}
            sols.append(cur)
            continue
        m = re_tour.match(line)
        if m and cur is not None:
            cur["tour"] = [int(v) for v in m.group(1).split(",")]
            cur["latency"] = int(m.group(2))
            continue
        m = re_init.search(line)
        if m and cur is not None:
            name, total = m.group(2), int(m.group(3))
            key = ("total" if total else "partial")
            kern = cur["kernels"].setdefault(name, {"nb": NB[name]}).setdefault(key, {})
            continue
        if kern is None:
            continue
        m = re_elems.search(line)
        if m:
            kern["move_elems"] = int(m.group(1))
            continue
        m = re_blck.search(line)
        if m:
            kern["grid_rows"] = int(m.group(1))
            kern["block"] = int(m.group(2))
            continue
        m = re_best.search(line)
        if m:
            kern["best"] = [int(m.group(1)), int(m.group(2)), int(m.group(3))]
            continue
        m = re_graph.search(line)
        if m:
            kern["graph"] = [int(m.group(1)), int(m.group(2))]
            continue
        m = re_imp.search(line)
        if m:
            kern["merge_" + m.group(2).lower()] = [int(m.group(3)), int(m.group(4))]
            continue
    return seed, sols


def main():
    out = []
    for fname in sorted(INST):
        seed, sols = parse(os.path.join(REF, fname))
        out.append({"log": fname, "instance": INST[fname], "seed": seed, "solutions": sols})
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "results_logs.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst, sum(len(o["solutions"]) for o in out), "solutions")


if __name__ == "__main__":
    main()
