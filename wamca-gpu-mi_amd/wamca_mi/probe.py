"""Per-launch timing of the sweep by tile shape and neighborhood mask.

``python -m wamca_mi.probe --n 1000 --tours 200`` builds one uniform instance
(nint distances) with ``tours`` random-permutation tours and, for every
(tile shape, mask) pair, times one launch over all tours and one launch over
the first tour (HIP events around each launch on the engine's stream).  It
prints one JSON line per launch in launch order, so counter or trace output
of the same run can be matched to them.
"""
from __future__ import annotations

import argparse
import json

from . import inputs as I
from .engine import Engine, rows_of


def moves(mask: int, N: int) -> int:
    per = {0: (N - 4) * (N - 3) // 2, 1: rows_of(1, 0, N) * (N - 3)}
    for k in (1, 2, 3):
        per[k + 1] = (N - k - 2) * (N - k - 1)
    return sum(v for nb, v in per.items() if (mask >> nb) & 1)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1000)
    ap.add_argument("--tours", type=int, default=200)
    ap.add_argument("--tiles", default="2,4", help="tile shape ids (engine TILES)")
    ap.add_argument("--masks", default="31,29,2,1,4")
    a = ap.parse_args(argv)
    eng = Engine(0)
    xs, ys = I.uniform(a.n, 1, 1000)
    inst = eng.instance(xs, ys, 1)
    slots = eng.slots(inst, a.tours)
    eng.set_tours(slots, [I.perm_tour(s, a.n) for s in range(1, a.tours + 1)])
    eng.sync()
    N = a.n + 1
    for tile in (int(t) for t in a.tiles.split(",")):
        eng.set_tile(tile)
        for mask in (int(m) for m in a.masks.split(",")):
            for mode, sl in (("batched", slots), ("single", slots[:1])):
                eng.sync()
                eng.region_begin()
                eng.eval(sl, mask, 0)
                us = eng.region_end() * 1e3
                mv = moves(mask, N) * len(sl)
                print(json.dumps({"mode": mode, "tile": tile, "mask": mask, "tours": len(sl),
                                  "moves": mv, "us": round(us, 2),
                                  "gmoves_per_s": round(mv / us / 1e3, 1)}), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
