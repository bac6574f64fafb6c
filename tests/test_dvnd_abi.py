"""The reference's extern "C" DVND interface (include/dvnd.h) served by
libwamca_mi.so, called by name through ctypes exactly as an external driver
would.  noConflict / getNoConflictMoves have no reference outputs: they are
checked against an independent restatement of dvnd.cu:85-175 and a truth table
of the reference's asymmetric cases.  applyMoves and bestNeighbor are checked
against the oracle.  The process-wide problem caches of the reference are kept
(first file wins), so bestNeighbor always uses berlin52 and applyMoves kroD100."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import GOLDEN, load_tsplib_coords
from oracle import oracle as O

wm = pytest.importorskip("wamca_mi")

BERLIN = os.path.join(GOLDEN, "instances", "01_berlin52.tsp").encode()
KROD = os.path.join(GOLDEN, "instances", "02_kroD100.tsp").encode()


@pytest.fixture(scope="module")
def L():
    lib = C.CDLL(wm.LIB_PATH)
    lib.noConflict.restype = C.c_bool
    lib.noConflict.argtypes = [C.c_ushort, C.c_uint, C.c_uint, C.c_ushort, C.c_uint, C.c_uint]
    lib.getNoConflictMoves.restype = C.c_int
    lib.getNoConflictMoves.argtypes = [C.c_uint, C.POINTER(C.c_ushort), C.POINTER(C.c_uint),
                                       C.POINTER(C.c_uint), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_bool, C.c_bool]
    lib.applyMoves.restype = C.c_uint
    lib.applyMoves.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.c_uint, C.c_uint,
                               C.POINTER(C.c_ushort), C.POINTER(C.c_uint), C.POINTER(C.c_uint),
                               C.POINTER(C.c_int)]
    lib.bestNeighbor.restype = C.c_uint
    lib.bestNeighbor.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.c_uint, C.c_int, C.c_bool,
                                 C.c_uint, C.POINTER(C.c_uint), C.POINTER(C.c_ushort),
                                 C.POINTER(C.c_uint), C.POINTER(C.c_uint), C.POINTER(C.c_int),
                                 C.c_bool, C.c_uint, C.POINTER(C.c_int)]
    lib.bestNeighborSimple.restype = C.c_uint
    lib.bestNeighborSimple.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.c_uint, C.c_int]
    return lib


# ---- independent restatement of dvnd.cu:132-175 -------------------------------
def ref_no_conflict(m1, m2):
    id1, i1, j1 = m1
    id2, i2, j2 = m2
    i1, j1, i2, j2 = i1 & 0x3FFF, j1 & 0x3FFF, i2 & 0x3FFF, j2 & 0x3FFF
    id1, id2 = id1 & 15, id2 & 15

    def kk(i):
        return 1 if i == 2 else (2 if i == 3 else 3)

    if id1 == 0:
        if id2 == 0:
            return abs(i1 - i2) > 1 and abs(i1 - j2) > 1 and abs(j1 - i2) > 1 and abs(j1 - j2) > 1
        if id2 == 1:
            return (i1 < i2 - 1 or i1 > j2 - 1) and (j1 < i2 - 1 or j1 > j2 + 1)
        k2 = kk(id2)
        return (j1 < min(i2, j2) - 1 or i1 > max(i2, j2) + k2 or
                (i1 < min(i2, j2) - 1 and j1 > max(i2, j2) + k2))
    if id1 == 1:
        if id2 == 0:
            return (i2 < i1 - 1 or i2 > j1 + 1) and (j2 < i1 - 1 or j2 > j1 + 1)
        if id2 == 1:
            return j1 < i2 - 1 or i1 > j2 + 1 or j2 > i1 - 1 or i2 > j1 + 1
        k2 = kk(id2)
        return i1 > max(i2, j2) + k2 or j1 < min(i2, j2) - 1
    k1 = kk(id1)
    if id2 == 0:
        return (j2 < min(i1, i2) - 1 or i2 > max(i1, i2) + k1 or
                (i2 < min(i1, i2) - 1 and j2 > max(i1, i2) + k1))
    if id2 == 1:
        return j2 < min(i1, j1) - 1 or i2 > max(i1, j1) + k1
    k2 = kk(id2)
    return max(i1, j1) + k1 < min(i2, j2) or min(i1, j1) > max(i2, j2) + k2


def test_no_conflict_matches_restatement(L):
    rng = np.random.RandomState(3)
    for _ in range(20000):
        a = (int(rng.randint(0, 5)), int(rng.randint(1, 60)), int(rng.randint(1, 60)))
        b = (int(rng.randint(0, 5)), int(rng.randint(1, 60)), int(rng.randint(1, 60)))
        assert bool(L.noConflict(*a, *b)) == ref_no_conflict(a, b), (a, b)


def test_no_conflict_reference_quirks(L):
    # swap x swap is the symmetric |delta| > 1 test
    assert L.noConflict(0, 5, 9, 0, 12, 20) and not L.noConflict(0, 5, 9, 0, 10, 20)
    # swap x 2-opt compares move1.i with move2.j - 1 (not + 1): i1 = j2 passes
    assert L.noConflict(0, 8, 30, 1, 3, 8) is True
    # 2-opt x 2-opt: the third clause is (move2.j > move1.i - 1), true for most pairs
    assert L.noConflict(1, 10, 20, 1, 12, 18) is True
    assert L.noConflict(1, 10, 20, 1, 1, 5) is True
    # the predicate is not symmetric
    cand = [(mid, i, j) for mid in range(5) for i in (2, 9, 17) for j in (5, 14, 30)]
    asym = [(a, b) for a in cand for b in cand if ref_no_conflict(a, b) != ref_no_conflict(b, a)]
    assert len(asym) > 50, "expected asymmetric cases"
    for a, b in asym:
        assert bool(L.noConflict(*a, *b)) == ref_no_conflict(a, b)
    # bit-field truncation: i = 16384 + 5 is stored as 5
    assert bool(L.noConflict(0, 16384 + 5, 9, 0, 12, 20)) == bool(L.noConflict(0, 5, 9, 0, 12, 20))


def ref_better_no_conflict(moves, maximize, melhor):
    """dvnd.cu:85-130 with distinct costs (so the sort order is unique)."""
    idx = list(range(len(moves)))
    desc = maximize if melhor else not maximize
    idx.sort(key=lambda q: moves[q][3], reverse=desc)
    alive = [True] * len(idx)
    for a in range(len(idx)):
        for b in range(a + 1, len(idx)):
            if not ref_no_conflict(moves[idx[a]][:3], moves[idx[b]][:3]):
                alive[b if melhor else a] = False
                break
    sel = [idx[a] for a in range(len(idx)) if alive[a]]
    return sel, sum(moves[q][3] for q in sel)


@pytest.mark.parametrize("maximize,melhor", [(False, False), (False, True), (True, False),
                                             (True, True)])
def test_get_no_conflict_moves(L, maximize, melhor):
    rng = np.random.RandomState(11)
    for trial in range(200):
        m = int(rng.randint(1, 25))
        costs = rng.permutation(np.arange(-5000, 5000))[:m]
        moves = [(int(rng.randint(0, 5)), int(rng.randint(1, 80)), int(rng.randint(1, 80)),
                  int(costs[q])) for q in range(m)]
        ids = (C.c_ushort * m)(*[x[0] for x in moves])
        iis = (C.c_uint * m)(*[x[1] for x in moves])
        jjs = (C.c_uint * m)(*[x[2] for x in moves])
        cs = (C.c_int * m)(*[x[3] for x in moves])
        sel = (C.c_int * m)()
        imp = (C.c_int * 1)()
        cnt = L.getNoConflictMoves(m, ids, iis, jjs, cs, sel, imp, maximize, melhor)
        rsel, rimp = ref_better_no_conflict(moves, maximize, melhor)
        assert list(sel[:cnt]) == rsel and imp[0] == rimp, trial


def test_apply_moves(L):
    xs, ys = load_tsplib_coords("02_kroD100.tsp")
    W = O.weights(xs, ys, 0)
    rng = O.MT(5)
    t0 = O.random_tour(rng, W, 0.5)
    N = len(t0)
    moves = [(0, 3, 40), (1, 50, 60), (2, 70, 75), (3, 20, 10), (4, 80, 85), (2, 30, 12)]
    ref = t0.copy()
    for mid, i, j in moves:
        nb = 0 if mid in (0, 1) else mid  # id 1 is applied as a swap (dvnd.cpp:51-75)
        ref = O.apply(nb, ref, i, j)
    sol = (C.c_int * N)(*t0.tolist())
    m = len(moves)
    ids = (C.c_ushort * m)(*[x[0] for x in moves])
    iis = (C.c_uint * m)(*[x[1] for x in moves])
    jjs = (C.c_uint * m)(*[x[2] for x in moves])
    cs = (C.c_int * m)(*([0] * m))
    val = L.applyMoves(KROD, sol, N, m, ids, iis, jjs, cs)
    assert list(sol) == ref.tolist()
    assert val == O.latency_u32(W, ref)


def test_best_neighbor_just_calc(L):
    xs, ys = load_tsplib_coords("01_berlin52.tsp")
    W = O.weights(xs, ys, 0)
    t = O.random_tour(O.MT(0), W, 0.5)
    sol = (C.c_int * len(t))(*t.tolist())
    assert L.bestNeighbor(BERLIN, sol, len(t), 0, True, 0, None, None, None, None, None,
                          False, 1, None) == 494309


@pytest.mark.gpu
@pytest.mark.parametrize("nb", [0, 1, 2, 3, 4])
def test_best_neighbor_gpu(L, nb):
    """One neighborhood step with the shipped ODC row partition: returned
    latency, new tour (solutionResp), and the recorded applied moves equal
    the oracle's merge of the same rows."""
    xs, ys = load_tsplib_coords("01_berlin52.tsp")
    W = O.weights(xs, ys, 0)
    t = O.random_tour(O.MT(7 + nb), W, 0.5)
    N = len(t)
    r = O.eval_rows(nb, 1, W, t)
    kept, s = O.merge(nb, r["cost"], r["i"], r["j"])
    t2, c, s2 = O.step(nb, 1, W, t)
    sol = (C.c_int * N)(*t.tolist())
    resp = (C.c_int * N)()
    cap = 64
    use = (C.c_uint * 1)(cap)
    ids = (C.c_ushort * cap)()
    iis = (C.c_uint * cap)()
    jjs = (C.c_uint * cap)()
    cs = (C.c_int * cap)()
    val = L.bestNeighbor(BERLIN, sol, N, nb, False, 0, use, ids, iis, jjs, cs, False, 1, resp)
    assert val == O.latency_u32(W, t2)
    assert list(resp) == t2.tolist()
    assert list(sol) == t.tolist()  # moves requested: the input tour is left alone
    assert use[0] == len(kept)
    for q, row in enumerate(kept):
        assert (ids[q], iis[q], jjs[q], cs[q]) == (nb, r["i"][row], r["j"][row], r["cost"][row])
    # without moves the new tour is written back into solution
    sol2 = (C.c_int * N)(*t.tolist())
    zero = (C.c_uint * 1)(0)
    val2 = L.bestNeighbor(BERLIN, sol2, N, nb, False, 0, zero, None, None, None, None, False, 1,
                          None)
    assert val2 == val and list(sol2) == t2.tolist()


@pytest.mark.gpu
def test_best_neighbor_berlin52_swap_pins_367403(L):
    xs, ys = load_tsplib_coords("01_berlin52.tsp")
    W = O.weights(xs, ys, 0)
    t = O.random_tour(O.MT(0), W, 0.5)
    sol = (C.c_int * len(t))(*t.tolist())
    zero = (C.c_uint * 1)(0)
    assert L.bestNeighbor(BERLIN, sol, len(t), 0, False, 0, zero, None, None, None, None,
                          False, 1, None) == 367403


GIL = os.path.join(GOLDEN, "instances", "09_gil262.tsp").encode()


def test_just_calc_negative_coordinates(L, monkeypatch):
    """gil262 (negative coordinates): the latency of a tour, from the raw
    TSPLIB coordinates as the reference reads them."""
    monkeypatch.setenv("MLP_DVND_STRICT", "0")  # problems keyed by hostCode
    xs, ys = load_tsplib_coords("09_gil262.tsp")
    W = O.weights(xs, ys, 0)
    t = O.random_tour(O.MT(1), W, 0.5)
    sol = (C.c_int * len(t))(*t.tolist())
    assert L.bestNeighbor(GIL, sol, len(t), 0, True, 262, None, None, None, None, None,
                          False, 1, None) == O.latency_u32(W, t)


@pytest.mark.gpu
@pytest.mark.parametrize("nb", [0, 1, 2, 3, 4])
def test_best_neighbor_negative_coordinates(L, nb, monkeypatch):
    monkeypatch.setenv("MLP_DVND_STRICT", "0")
    xs, ys = load_tsplib_coords("09_gil262.tsp")
    W = O.weights(xs, ys, 0)
    t = O.random_tour(O.MT(20 + nb), W, 0.5)
    N = len(t)
    t2, c, _ = O.step(nb, 1, W, t)
    assert c > 0
    sol = (C.c_int * N)(*t.tolist())
    zero = (C.c_uint * 1)(0)
    val = L.bestNeighbor(GIL, sol, N, nb, False, 262, zero, None, None, None, None, False, 1,
                         None)
    assert val == O.latency_u32(W, t2)
    assert list(sol) == t2.tolist()
