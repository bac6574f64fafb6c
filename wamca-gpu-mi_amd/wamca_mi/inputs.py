"""Benchmark and test inputs produced by the library's host code (mlp_host.h):
synthetic instances, TSPLIB files and seeded start tours."""
from __future__ import annotations

import numpy as np

from ._lib import lib


def _ck(st, what):
    if st != 0:
        raise ValueError(f"{what} failed with status {st}")


def uniform(n: int, seed: int, maxc: int = 1000):
    xs = np.empty(n, np.int32)
    ys = np.empty(n, np.int32)
    _ck(lib().mlp_gen_uniform(n, seed, maxc, xs, ys), "uniform instance")
    return xs, ys


def clustered(n: int, seed: int, ncent: int = 20, span: int = 2000, sigma: float = 40.0):
    xs = np.empty(n, np.int32)
    ys = np.empty(n, np.int32)
    _ck(lib().mlp_gen_clustered(n, seed, ncent, span, sigma, xs, ys), "clustered instance")
    return xs, ys


def loguniform_n(seed: int, index: int, lo: int = 50, hi: int = 2000) -> int:
    return int(lib().mlp_gen_loguniform_n(seed, index, lo, hi))


def tsplib(path: str):
    cap = 1 << 16
    xs = np.empty(cap, np.int32)
    ys = np.empty(cap, np.int32)
    import ctypes as C
    n = C.c_int32()
    _ck(lib().mlp_tsplib_load(path.encode(), xs, ys, cap, C.byref(n)), f"TSPLIB {path}")
    return xs[:n.value].copy(), ys[:n.value].copy()


def perm_tour(seed: int, n: int):
    t = np.empty(n + 1, np.int32)
    _ck(lib().mlp_tour_perm(seed, n, t), "perm tour")
    return t


def random_tour(xs, ys, rnd: int, seed: int, skip: int = 0, alpha: float = 0.5):
    xs = np.ascontiguousarray(xs, np.int32)
    ys = np.ascontiguousarray(ys, np.int32)
    t = np.empty(len(xs) + 1, np.int32)
    _ck(lib().mlp_tour_random(xs, ys, len(xs), rnd, seed, skip, alpha, t), "random tour")
    return t


def random_tour_fast(xs, ys, rnd: int, seed: int, alpha: float = 0.5):
    """The randomized construction with selection instead of sorting (O(n^2);
    ties in distance ordered by client id) for large instances."""
    xs = np.ascontiguousarray(xs, np.int32)
    ys = np.ascontiguousarray(ys, np.int32)
    t = np.empty(len(xs) + 1, np.int32)
    _ck(lib().mlp_tour_random_fast(xs, ys, len(xs), rnd, seed, alpha, t), "random tour (fast)")
    return t
