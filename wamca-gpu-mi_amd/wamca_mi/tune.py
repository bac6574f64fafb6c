"""Sweep tile-shape autotuning.

``python -m wamca_mi.tune --out FILE`` times the five sweep tile shapes on
launches of every neighborhood mask a DVND round or a full sweep issues, over
instance sizes and tours per launch spanning 2^10 .. 2^33 moves, and writes the
fastest shape per (mask, log2 moves) as JSON.  ``Engine`` loads the table of
its architecture from ``tune_<arch>.json`` beside this file.
"""
from __future__ import annotations

import argparse
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MASKS = (1, 2, 4, 8, 16, 31)


def table_path(arch: str) -> str:
    return os.path.join(HERE, f"tune_{arch.split(':')[0]}.json")


def load_table(arch: str):
    try:
        with open(table_path(arch)) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    t = np.full((32, 48), -1, np.int8)
    for mask, row in d["table"].items():
        for b, tile in row.items():
            t[int(mask), int(b)] = int(tile)
    return t


def run(eng, sizes=(32, 64, 128, 256, 512, 1024, 2048, 4096, 8192),
        counts=(1, 2, 4, 8, 16, 64, 256), max_moves=1 << 33, reps=3, log=None):
    from . import inputs as I
    records = []
    for n in sizes:
        xs, ys = I.uniform(n, 11, 1000)
        inst = eng.instance(xs, ys, 1)
        N = n + 1
        per_tour = (N * N * 4)  # upper bound of the five neighborhoods' moves
        cs = [c for c in counts if c * per_tour <= max_moves]
        if not cs:
            eng.destroy_instance(inst)
            continue
        slots = eng.slots(inst, max(cs))
        eng.set_tours(slots, [I.perm_tour(q + 1, n) for q in range(len(slots))])
        for c in cs:
            for mask in MASKS:
                best, us = eng.autotune(slots[:c], mask, reps)
                records.append({"n": n, "count": c, "mask": mask, "best": best, "us": us})
                if log:
                    log(f"n={n} count={c} mask={mask} best={best} us={[round(u, 1) for u in us]}")
        for s in slots:
            eng.destroy_slot(s)
        eng.destroy_instance(inst)
    return records


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    from .engine import Engine
    eng = Engine(0, tuned=False)
    arch = eng.device_info()["arch"]
    recs = run(eng, log=lambda m: print(m, flush=True))
    t = eng.tune_table()
    table = {str(m): {str(b): int(t[m, b]) for b in range(48) if t[m, b] >= 0}
             for m in range(32) if (t[m] >= 0).any()}
    out = args.out or table_path(arch)
    with open(out, "w") as f:
        json.dump({"arch": arch, "table": table, "measurements": recs}, f, indent=0)
    print(f"wrote {out}")
    eng.close()


if __name__ == "__main__":
    main()
