"""ctypes view of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg import
this module.  It loads ``oracle/_build/libmlp_oracle.so`` (built by
``oracle/Makefile``) and exposes numpy-friendly wrappers around the plain
restatement of the reference's MLP local search in ``mlp_oracle.cpp``.

Conventions: a tour is an int32 array of N = n + 1 client ids that starts and
ends with the depot 0; ``W`` is the client-indexed n x n int32 weight matrix;
``nb`` 0 = swap, 1 = 2-opt, 2..4 = or-opt-1..3; ``rowmap`` 0 = ADS kernels'
row partition, 1 = ODC kernels' row partition.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libmlp_oracle.so")
_lib = None

I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
I64P = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
U32P = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        sig = {
            "or_mt_state_bytes": (C.c_int, []),
            "or_mt_seed": (None, [vp, C.c_uint32]),
            "or_mt_rand": (C.c_uint32, [vp]),
            "or_mt_randmax": (C.c_uint32, [vp, C.c_uint32]),
            "or_mt_real_oo": (C.c_double, [vp]),
            "or_mt_real_co": (C.c_double, [vp]),
            "or_mt_shuffle_i32": (None, [vp, I32P, C.c_uint32]),
            "or_dist": (C.c_int, [C.c_int] * 4 + [C.c_int]),
            "or_dist_s": (C.c_int, [C.c_int32, C.c_int]),
            "or_weights": (None, [I32P, I32P, C.c_int, C.c_int, I32P]),
            "or_instance_uniform": (None, [C.c_int, C.c_uint32, C.c_int, I32P, I32P]),
            "or_instance_clustered": (None, [C.c_int, C.c_uint32, C.c_int, C.c_int, C.c_double,
                                             I32P, I32P]),
            "or_random_tour": (None, [vp, I32P, C.c_int, C.c_float, I32P]),
            "or_perm_tour": (None, [C.c_uint32, C.c_int, I32P]),
            "or_latency": (C.c_int64, [I32P, C.c_int, I32P, C.c_int]),
            "or_latency_u32": (C.c_uint32, [I32P, C.c_int, I32P, C.c_int]),
            "or_ads_u32": (None, [I32P, C.c_int, I32P, C.c_int, U32P, U32P]),
            "or_rows": (C.c_int, [C.c_int, C.c_int, C.c_int]),
            "or_ctxs": (C.c_int, [C.c_int, C.c_int, C.c_int]),
            "or_delta_ads": (C.c_int64, [C.c_int, I32P, C.c_int, I32P, C.c_int, C.c_int, C.c_int]),
            "or_delta_odc": (C.c_int64, [C.c_int, I32P, C.c_int, I32P, C.c_int, C.c_int, C.c_int]),
            "or_delta_brute": (C.c_int64, [C.c_int, I32P, C.c_int, I32P, C.c_int, C.c_int,
                                           C.c_int]),
            "or_enumerate": (C.c_int64, [C.c_int, C.c_int, I32P, C.c_int, I32P, C.c_int,
                                         vp, vp, vp, vp, vp, vp, vp]),
            "or_eval_rows": (C.c_int, [C.c_int, C.c_int, I32P, C.c_int, I32P, C.c_int,
                                       I64P, I32P, I32P, I32P]),
            "or_eval_total": (C.c_int64, [C.c_int, I32P, C.c_int, I32P, C.c_int, C.c_int,
                                          C.c_void_p, C.c_void_p, C.c_void_p]),
            "or_merge": (C.c_int, [C.c_int, C.c_int, I64P, I32P, I32P, I32P, I64P]),
            "or_apply": (None, [C.c_int, I32P, C.c_int, C.c_int]),
            "or_can_merge": (C.c_int, [C.c_int] * 5),
            "or_step": (C.c_int, [C.c_int, C.c_int, I32P, C.c_int, I32P, C.c_int, I64P]),
            "or_dvnd": (C.c_int64, [C.c_int, I32P, C.c_int, I32P, C.c_int, C.c_int, I32P, I64P]),
            "or_dvnd_nbs": (C.c_int64, [C.c_int, I32P, C.c_int, I32P, C.c_int, C.c_int, C.c_int,
                                        I32P, I64P]),
            "or_delta_prefix": (C.c_int64, [C.c_int, I32P, I32P, C.c_int, I32P, C.c_int, C.c_int,
                                            C.c_int]),
            "or_latency_xy": (C.c_int64, [I32P, I32P, C.c_int, I32P, C.c_int]),
            "or_eval_rows_xy": (C.c_int, [C.c_int, C.c_int, I32P, I32P, C.c_int, I32P, C.c_int,
                                          I64P, I32P, I32P, I32P, C.c_int]),
            "or_dvnd_xy": (C.c_int64, [C.c_int, I32P, I32P, C.c_int, I32P, C.c_int, C.c_int,
                                       C.c_int, I32P, I64P, C.c_int]),
            "or_eval_batch_xy": (C.c_int64, [C.c_uint32, C.c_int, I32P, I32P, C.c_int, I32P,
                                             C.c_int, C.c_int, I64P, I32P, I32P, C.c_int]),
            "or_max_threads": (C.c_int, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class MT:
    """MT19937 state (mtrand.cpp:51-121)."""

    def __init__(self, seed: int):
        self._buf = C.create_string_buffer(lib().or_mt_state_bytes())
        lib().or_mt_seed(self._buf, seed)

    def rand(self) -> int:
        return lib().or_mt_rand(self._buf)

    def randmax(self, m: int) -> int:
        return lib().or_mt_randmax(self._buf, m)

    def real_oo(self) -> float:
        return lib().or_mt_real_oo(self._buf)

    def real_co(self) -> float:
        return lib().or_mt_real_co(self._buf)

    @property
    def ptr(self):
        return self._buf


def dist(x1, y1, x2, y2, rnd=0) -> int:
    return lib().or_dist(x1, y1, x2, y2, rnd)


def weights(xs, ys, rnd=0) -> np.ndarray:
    xs = np.ascontiguousarray(xs, np.int32)
    ys = np.ascontiguousarray(ys, np.int32)
    n = len(xs)
    W = np.empty(n * n, np.int32)
    lib().or_weights(xs, ys, n, rnd, W)
    return W.reshape(n, n)


def instance_uniform(n, seed, maxc=1000):
    xs = np.empty(n, np.int32)
    ys = np.empty(n, np.int32)
    lib().or_instance_uniform(n, seed, maxc, xs, ys)
    return xs, ys


def instance_clustered(n, seed, ncent=20, span=2000, sigma=40.0):
    xs = np.empty(n, np.int32)
    ys = np.empty(n, np.int32)
    lib().or_instance_clustered(n, seed, ncent, span, sigma, xs, ys)
    return xs, ys


def random_tour(rng: MT, W: np.ndarray, alpha=0.5) -> np.ndarray:
    n = W.shape[0]
    t = np.empty(n + 1, np.int32)
    lib().or_random_tour(rng.ptr, np.ascontiguousarray(W, np.int32).ravel(), n, alpha, t)
    return t


def perm_tour(seed: int, n: int) -> np.ndarray:
    t = np.empty(n + 1, np.int32)
    lib().or_perm_tour(seed, n, t)
    return t


def _wt(W, tour):
    W = np.ascontiguousarray(W, np.int32)
    return W.ravel(), W.shape[0], np.ascontiguousarray(tour, np.int32)


def latency(W, tour) -> int:
    Wf, n, t = _wt(W, tour)
    return lib().or_latency(Wf, n, t, len(t))


def latency_u32(W, tour) -> int:
    Wf, n, t = _wt(W, tour)
    return lib().or_latency_u32(Wf, n, t, len(t))


def ads_u32(W, tour):
    Wf, n, t = _wt(W, tour)
    N = len(t)
    T = np.empty(N, np.uint32)
    Cm = np.empty(N * N, np.uint32)
    lib().or_ads_u32(Wf, n, t, N, T, Cm)
    return T, Cm.reshape(N, N)


def rows(nb, rowmap, N) -> int:
    return lib().or_rows(nb, rowmap, N)


def delta(kind, nb, W, tour, i, j) -> int:
    Wf, n, t = _wt(W, tour)
    f = {"ads": lib().or_delta_ads, "odc": lib().or_delta_odc, "brute": lib().or_delta_brute}[kind]
    return f(nb, Wf, n, t, len(t), i, j)


def enumerate_moves(nb, rowmap, W, tour, ads=True, odc=False, brute=False):
    """All moves of a row map in (row, ctx) order with the requested costs."""
    Wf, n, t = _wt(W, tour)
    N = len(t)
    L = lib()
    m = L.or_enumerate(nb, rowmap, Wf, n, t, N, None, None, None, None, None, None, None)
    out = {k: np.empty(m, np.int32) for k in ("row", "ctx", "i", "j")}
    d = {k: (np.empty(m, np.int64) if on else None)
         for k, on in (("ads", ads), ("odc", odc), ("brute", brute))}

    def p(a):
        return a.ctypes.data_as(C.c_void_p) if a is not None else None

    L.or_enumerate(nb, rowmap, Wf, n, t, N, p(out["row"]), p(out["ctx"]), p(out["i"]),
                   p(out["j"]), p(d["ads"]), p(d["odc"]), p(d["brute"]))
    out.update({k: v for k, v in d.items() if v is not None})
    return out


def eval_rows(nb, rowmap, W, tour):
    """Per-row best records: dict(cost int64, i, j, ctx) of length rows."""
    Wf, n, t = _wt(W, tour)
    N = len(t)
    R = max(rows(nb, rowmap, N), 0)
    cost = np.empty(R, np.int64)
    i = np.empty(R, np.int32)
    j = np.empty(R, np.int32)
    ctx = np.empty(R, np.int32)
    lib().or_eval_rows(nb, rowmap, Wf, n, t, N, cost, i, j, ctx)
    return {"cost": cost, "i": i, "j": j, "ctx": ctx}


def eval_total(nb, W, tour, block):
    """TOTAL-mode entries (ADS row map): row * block + t holds the first
    strict minimum over the row-local indices t, t + block, ...; dict(cost
    int64 (INT64_MAX when empty), i, j) of length rows * block."""
    Wf, n, t = _wt(W, tour)
    N = len(t)
    M = lib().or_eval_total(nb, Wf, n, t, N, block, None, None, None)
    cost = np.empty(M, np.int64)
    i = np.empty(M, np.int32)
    j = np.empty(M, np.int32)
    if M:
        lib().or_eval_total(nb, Wf, n, t, N, block, cost.ctypes.data, i.ctypes.data,
                            j.ctypes.data)
    return {"cost": cost, "i": i, "j": j}


def merge(nb, cost, i, j):
    """Greedy merge of a per-row list: (kept row ids in sorted order, sum)."""
    cost = np.ascontiguousarray(cost, np.int64)
    i = np.ascontiguousarray(i, np.int32)
    j = np.ascontiguousarray(j, np.int32)
    kept = np.empty(len(cost), np.int32)
    s = np.zeros(1, np.int64)
    c = lib().or_merge(nb, len(cost), cost, i, j, kept, s)
    return kept[:c].copy(), int(s[0])


def apply(nb, tour, i, j) -> np.ndarray:
    t = np.ascontiguousarray(tour, np.int32).copy()
    lib().or_apply(nb, t, i, j)
    return t


def can_merge(nb, i1, j1, i2, j2) -> bool:
    return bool(lib().or_can_merge(nb, i1, j1, i2, j2))


def step(nb, rowmap, W, tour):
    """Evaluate + merge + apply one neighborhood: (new tour, count, sum)."""
    Wf, n, t = _wt(W, tour)
    t = t.copy()
    s = np.zeros(1, np.int64)
    c = lib().or_step(nb, rowmap, Wf, n, t, len(t), s)
    return t, c, int(s[0])


def dvnd(W, tour, rowmap=0, max_rounds=0, nbs=5):
    """DVND over neighborhood ids 0..nbs-1 to a local optimum:
    (final tour, latency, rounds, moves)."""
    Wf, n, t = _wt(W, tour)
    t = t.copy()
    r = np.zeros(1, np.int32)
    m = np.zeros(1, np.int64)
    L = lib().or_dvnd_nbs(rowmap, Wf, n, t, len(t), max_rounds, nbs, r, m)
    return t, int(L), int(r[0]), int(m[0])


# -- O(N)-memory forms over coordinates (the ADS in prefix form) ---------------
def _xy(xs, ys, tour):
    return (np.ascontiguousarray(xs, np.int32), np.ascontiguousarray(ys, np.int32),
            np.ascontiguousarray(tour, np.int32))


def delta_prefix(nb, xs, ys, rnd, tour, i, j) -> int:
    """Move delta from the prefix form of the ADS and coordinate distances."""
    xs, ys, t = _xy(xs, ys, tour)
    return lib().or_delta_prefix(nb, xs, ys, rnd, t, len(t), i, j)


def latency_xy(xs, ys, rnd, tour) -> int:
    xs, ys, t = _xy(xs, ys, tour)
    return lib().or_latency_xy(xs, ys, rnd, t, len(t))


def eval_rows_xy(nb, rowmap, xs, ys, rnd, tour, threads=0):
    """eval_rows over coordinates (rows split over threads)."""
    xs, ys, t = _xy(xs, ys, tour)
    N = len(t)
    R = max(rows(nb, rowmap, N), 0)
    cost = np.empty(R, np.int64)
    i = np.empty(R, np.int32)
    j = np.empty(R, np.int32)
    ctx = np.empty(R, np.int32)
    lib().or_eval_rows_xy(nb, rowmap, xs, ys, rnd, t, N, cost, i, j, ctx, threads)
    return {"cost": cost, "i": i, "j": j, "ctx": ctx}


def dvnd_xy(xs, ys, rnd, tour, rowmap=0, max_rounds=0, nbs=5, threads=0):
    """dvnd over coordinates in O(N) memory: (final tour, latency, rounds, moves)."""
    xs, ys, t = _xy(xs, ys, tour)
    t = t.copy()
    r = np.zeros(1, np.int32)
    m = np.zeros(1, np.int64)
    L = lib().or_dvnd_xy(rowmap, xs, ys, rnd, t, len(t), max_rounds, nbs, r, m, threads)
    return t, int(L), int(r[0]), int(m[0])


def eval_batch_xy(nb_mask, rowmap, xs, ys, rnd, tours, threads=0):
    """Per-row bests of the neighborhoods in nb_mask for every tour (OpenMP
    over tour x row).  Returns ({nb: {"cost", "i", "j"} arrays of shape
    (tours, rows)}, moves evaluated)."""
    xs = np.ascontiguousarray(xs, np.int32)
    ys = np.ascontiguousarray(ys, np.int32)
    flat = np.ascontiguousarray(np.stack([np.asarray(t, np.int32) for t in tours]))
    count, N = flat.shape
    nbs = [nb for nb in range(5) if (nb_mask >> nb) & 1]
    R = {nb: max(rows(nb, rowmap, N), 0) for nb in nbs}
    stride = sum(R.values())
    cost = np.empty(count * stride, np.int64)
    i = np.empty(count * stride, np.int32)
    j = np.empty(count * stride, np.int32)
    moves = lib().or_eval_batch_xy(nb_mask, rowmap, xs, ys, rnd, flat.ravel(), count, N, cost, i,
                                   j, threads)
    out, off = {}, 0
    for nb in nbs:
        sl = slice(off, off + R[nb])
        out[nb] = {k: a.reshape(count, stride)[:, sl] for k, a in (("cost", cost), ("i", i),
                                                                   ("j", j))}
        off += R[nb]
    return out, int(moves)


def max_threads() -> int:
    return int(lib().or_max_threads())
