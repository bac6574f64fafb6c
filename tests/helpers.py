"""Shared fixtures for the tests: TSPLIB coordinates and the log golden file."""
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_tsplib_coords(name):
    """EUC_2D coordinates as the reference reads them: parsed into float
    (tsplib.h:56 ``typedef float Float``) and rounded with inear
    (tsplib.cpp:236-240), node ids in file order (mlproblem.cpp:110-117)."""
    xs, ys, sec = [], [], False
    half = np.float32(0.5)

    def inear(v):
        v = np.float32(v)
        return int(np.float32(v + half)) if v >= 0 else int(np.float32(v - half))

    with open(os.path.join(GOLDEN, "instances", name)) as f:
        for line in f:
            tok = line.split()
            if not tok:
                continue
            if tok[0].startswith("NODE_COORD_SECTION"):
                sec = True
                continue
            if tok[0] == "EOF":
                break
            if sec:
                xs.append(inear(float(tok[1])))
                ys.append(inear(float(tok[2])))
    return np.array(xs, np.int32), np.array(ys, np.int32)


def results_logs():
    with open(os.path.join(GOLDEN, "results_logs.json")) as f:
        return json.load(f)
