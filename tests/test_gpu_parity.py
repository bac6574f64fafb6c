"""HIP path vs the CPU oracle, bit for bit, through the C ABI (libwamca_mi.so).

Distances are checked exhaustively for every squared length s < 2^24 and on
the rounding edges above; per-row best moves, merged move sets, applied tours
and DVND fixed points are compared with the oracle on the reference's own
instances, the logged start tours and seeded synthetic tours.
"""
import numpy as np
import pytest

from helpers import load_tsplib_coords, results_logs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

wm = pytest.importorskip("wamca_mi")


@pytest.fixture(scope="module")
def eng():
    e = wm.Engine(0)
    yield e
    e.close()


def np_dist_s(s, rnd):
    f = np.asarray(s, np.int64).astype(np.float32)
    r = np.sqrt(f)
    return (r + np.float32(0.5 if rnd else 0.0)).astype(np.int32)


@pytest.mark.parametrize("rnd", [0, 1])
def test_distance_exhaustive_below_2p24(eng, rnd):
    step = 1 << 22
    for s0 in range(0, 1 << 24, step):
        got = eng.dist_table(s0, step, rnd)
        ref = np_dist_s(np.arange(s0, s0 + step, dtype=np.int64), rnd)
        bad = np.nonzero(got != ref)[0]
        assert bad.size == 0, (rnd, s0 + int(bad[0]), int(got[bad[0]]), int(ref[bad[0]]))


@pytest.mark.parametrize("rnd", [0, 1])
def test_distance_rounding_edges(eng, rnd):
    """Squared lengths next to k^2 and k^2 + k (where truncation and rounding
    switch) for every k up to 46340, i.e. up to the int32 limit of dx^2+dy^2."""
    k = np.arange(1, 46341, dtype=np.int64)
    cand = np.concatenate([k * k - 1, k * k, k * k + 1, k * k + k - 1, k * k + k, k * k + k + 1])
    cand = np.unique(cand[(cand >= 0) & (cand <= 2**31 - 1)]).astype(np.int32)
    got = eng.dist_list(cand, rnd)
    ref = np_dist_s(cand, rnd)
    bad = np.nonzero(got != ref)[0]
    assert bad.size == 0, (int(cand[bad[0]]), int(got[bad[0]]), int(ref[bad[0]]))


def _inst(name):
    xs, ys = load_tsplib_coords(name)
    return xs, ys


@pytest.mark.parametrize("name,rnd", [("02_kroD100.tsp", 0), ("03_pr226.tsp", 0),
                                      ("05_TRP-S500-R1.tsp", 1)])
def test_weights(eng, name, rnd):
    xs, ys = _inst(name)
    inst = eng.instance(xs, ys, rnd)
    assert np.array_equal(eng.weights(inst), O.weights(xs, ys, rnd))


def _check_rows(eng, s, W, tour, nb, rowmap):
    got = eng.rows(s, nb)
    ref = O.eval_rows(nb, rowmap, W, tour)
    assert len(got["cost"]) == len(ref["cost"])
    empty = ref["ctx"] < 0
    rc = np.where(empty, np.iinfo(np.int64).max, ref["cost"])
    bad = np.nonzero((got["cost"] != rc) | ((got["i"] != ref["i"]) & ~empty) |
                     ((got["j"] != ref["j"]) & ~empty))[0]
    assert bad.size == 0, (nb, rowmap, int(bad[0]), got["cost"][bad[0]], rc[bad[0]],
                           (got["i"][bad[0]], got["j"][bad[0]]), (ref["i"][bad[0]], ref["j"][bad[0]]))


def test_logged_tours_rows_and_latency(eng):
    for log in results_logs():
        xs, ys = _inst(log["instance"])
        W = O.weights(xs, ys, 0)
        inst = eng.instance(xs, ys, 0)
        for sol in log["solutions"]:
            t = np.array(sol["tour"], np.int32)
            s = eng.slot(inst)
            eng.set_tour(s, t)
            assert int(eng.latency([s])[0]) % 2**32 == sol["latency"]
            eng.eval([s], 0x1F, 0)
            for name, k in sol["kernels"].items():
                p = k.get("partial", {})
                if "best" not in p:
                    continue
                r = eng.rows(s, k["nb"])
                b = int(np.argsort(r["cost"], kind="stable")[0])
                assert [int(r["i"][b]), int(r["j"][b]), int(r["cost"][b])] == p["best"]
                assert [int((r["cost"] < 0).sum()), len(r["cost"])] == p["graph"]
            for nb in range(5):
                _check_rows(eng, s, W, t, nb, 0)
            eng.destroy_slot(s)


@pytest.mark.parametrize("n,rnd,seed", [(52, 0, 3), (100, 1, 1), (226, 1, 2), (500, 1, 3),
                                        (1000, 1, 4), (2000, 1, 5)])
@pytest.mark.parametrize("rowmap", [0, 1])
def test_rows_random_tours(eng, n, rnd, seed, rowmap):
    xs, ys = O.instance_uniform(n, seed)
    W = O.weights(xs, ys, rnd)
    inst = eng.instance(xs, ys, rnd)
    s = eng.slot(inst)
    for tseed in (seed, seed + 100):
        t = O.perm_tour(tseed, n)
        eng.set_tour(s, t)
        assert int(eng.latency([s])[0]) == O.latency(W, t)
        eng.eval([s], 0x1F, rowmap)
        for nb in range(5):
            _check_rows(eng, s, W, t, nb, rowmap)


def test_clustered_and_tiny(eng):
    for n in (5, 6, 7, 8, 13):
        xs, ys = O.instance_uniform(n, n, 50)
        W = O.weights(xs, ys, 0)
        inst = eng.instance(xs, ys, 0)
        s = eng.slot(inst)
        t = O.perm_tour(n, n)
        eng.set_tour(s, t)
        for rowmap in (0, 1):
            eng.eval([s], 0x1F, rowmap)
            for nb in range(5):
                if O.rows(nb, rowmap, n + 1) > 0:
                    _check_rows(eng, s, W, t, nb, rowmap)
    xs, ys = O.instance_clustered(784, 9)
    W = O.weights(xs, ys, 1)
    inst = eng.instance(xs, ys, 1)
    s = eng.slot(inst)
    t = O.random_tour(O.MT(9), W, 0.5)
    eng.set_tour(s, t)
    eng.eval([s], 0x1F, 0)
    for nb in range(5):
        _check_rows(eng, s, W, t, nb, 0)


@pytest.mark.parametrize("rowmap", [0, 1])
@pytest.mark.parametrize("tile", [-1, 0, 2])
@pytest.mark.parametrize("persistent", [True, False])
def test_persistent_and_batched_equal_single(eng, rowmap, tile, persistent):
    # 40 slots of one instance: more segments than a launch carries inline,
    # all of one size (one segment per slot with 2-opt in the fixed-row tiles
    # of 12x1024, two with 2-opt on its own tiles of 4x256 or under ODC)
    xs, ys = O.instance_uniform(300, 11)
    W = O.weights(xs, ys, 1)
    inst = eng.instance(xs, ys, 1)
    S = eng.slots(inst, 40)
    tours = [O.perm_tour(500 + q, 300) for q in range(40)]
    eng.set_tours(S, tours)
    eng.set_tile(tile)
    try:
        eng.eval(S, 0x1F, rowmap, persistent=persistent)
        batched = {(s, nb): eng.rows(s, nb)["cost"] for s in S for nb in range(5)}
        for q, s in enumerate(S):
            eng.eval([s], 0x1F, rowmap)
            for nb in range(5):
                assert np.array_equal(eng.rows(s, nb)["cost"], batched[(s, nb)]), (q, nb)
    finally:
        eng.set_tile(-1)
    for q in (0, 7, 39):
        for nb in range(5):
            _check_rows(eng, S[q], W, tours[q], nb, rowmap)
    for s in S:
        eng.destroy_slot(s)
    eng.destroy_instance(inst)


def test_ads_materialized(eng):
    xs, ys = _inst("02_kroD100.tsp")
    W = O.weights(xs, ys, 0)
    inst = eng.instance(xs, ys, 0)
    s = eng.slot(inst)
    t = O.random_tour(O.MT(100), W, 0.5)
    eng.set_tour(s, t)
    buf = eng.ads(s)
    N = len(t)
    T, Cm = O.ads_u32(W, t)
    assert buf[0] == 128 + 4 * N * (N + 3) and buf[1] == N and buf[2] == N
    assert buf[3] == O.latency_u32(W, t) and buf[4] == 1 and buf[5] == 0
    coords = buf[32:32 + N]
    assert np.array_equal(coords >> 16, xs[t]) and np.array_equal(coords & 0xFFFF, ys[t])
    sol = buf[32 + N:32 + 2 * N]
    assert np.array_equal(sol[1:] >> 16, t[1:])
    assert np.array_equal(buf[32 + 2 * N:32 + 3 * N], T)
    assert np.array_equal(buf[32 + 3 * N:].reshape(N, N), Cm)


def _check_step(eng, s, W, tour, nb, rowmap):
    eng.set_tour(s, tour)
    eng.eval([s], 1 << nb, rowmap)
    eng.merge_apply([s], nb)
    got_moves = eng.applied(s, nb)
    r = O.eval_rows(nb, rowmap, W, tour)
    kept, ssum = O.merge(nb, r["cost"], r["i"], r["j"])
    ref_moves = [(int(r["cost"][q]), nb, int(r["i"][q]), int(r["j"][q]), int(q)) for q in kept]
    assert [tuple(int(v) for v in m) for m in got_moves] == ref_moves
    t2, c, s2 = O.step(nb, rowmap, W, tour)
    assert c == len(kept) and s2 == ssum
    assert np.array_equal(eng.get_tour(s), t2)
    assert int(eng.latency([s])[0]) == O.latency(W, t2)
    return t2


def test_berlin52_odc_swap_pass_gpu(eng):
    xs, ys = _inst("01_berlin52.tsp")
    W = O.weights(xs, ys, 0)
    inst = eng.instance(xs, ys, 0)
    s = eng.slot(inst)
    t = O.random_tour(O.MT(0), W, 0.5)
    eng.set_tour(s, t)
    assert int(eng.latency([s])[0]) == 494309
    eng.eval([s], 1, 1)
    eng.merge_apply([s], 0)
    assert int(eng.latency([s])[0]) == 367403
    _check_step(eng, s, W, t, 0, 1)


def test_logged_merges_gpu(eng):
    for log in results_logs():
        xs, ys = _inst(log["instance"])
        W = O.weights(xs, ys, 0)
        inst = eng.instance(xs, ys, 0)
        s = eng.slot(inst)
        for sol in log["solutions"]:
            t = np.array(sol["tour"], np.int32)
            for name, k in sol["kernels"].items():
                p = k.get("partial", {})
                if "merge_gpu" not in p:
                    continue
                eng.set_tour(s, t)
                eng.eval([s], 1 << k["nb"], 0)
                eng.merge_apply([s], k["nb"])
                mv = eng.applied(s, k["nb"])
                assert [sum(m[0] for m in mv), len(mv)] == p["merge_gpu"]
                assert int(eng.latency([s])[0]) == sol["latency"] + p["merge_gpu"][0]


@pytest.mark.parametrize("n,seed", [(100, 1), (300, 2), (1000, 3)])
@pytest.mark.parametrize("rowmap", [0, 1])
def test_merge_apply_steps(eng, n, seed, rowmap):
    xs, ys = O.instance_uniform(n, seed)
    W = O.weights(xs, ys, 1)
    inst = eng.instance(xs, ys, 1)
    s = eng.slot(inst)
    t = O.random_tour(O.MT(seed), W, 0.5)
    for nb in range(5):
        t = _check_step(eng, s, W, t, nb, rowmap)
    tp = O.perm_tour(seed, n)
    for nb in range(5):
        tp = _check_step(eng, s, W, tp, nb, rowmap)


def test_dvnd_n100_uniform(eng):
    """n = 100 uniform [0,1000]^2 seed 1, nint, random-insertion start (seed 1)."""
    xs, ys = O.instance_uniform(100, 1)
    W = O.weights(xs, ys, 1)
    t0 = O.random_tour(O.MT(1), W, 0.5)
    tref, Lref, rref, _ = O.dvnd(W, t0, 0)
    inst = eng.instance(xs, ys, 1)
    s = eng.slot(inst)
    eng.set_tour(s, t0)
    rounds = eng.dvnd([s], 0)
    assert int(eng.latency([s])[0]) == Lref
    assert np.array_equal(eng.get_tour(s), tref)
    assert int(rounds[0]) == rref


def test_dvnd_batch_uneven_equals_alone(eng):
    insts, tours, refs = [], [], []
    for q, n in enumerate([50, 77, 130, 260, 400]):
        if q % 2:
            xs, ys = O.instance_clustered(n, 40 + q)
        else:
            xs, ys = O.instance_uniform(n, 40 + q)
        W = O.weights(xs, ys, 1)
        t0 = O.random_tour(O.MT(40 + q), W, 0.5)
        insts.append(eng.instance(xs, ys, 1))
        tours.append(t0)
        refs.append(O.dvnd(W, t0, 0))
    S = [eng.slot(i) for i in insts]
    eng.set_tours(S, tours)
    rounds = eng.dvnd(S, 0)
    for q, s in enumerate(S):
        assert int(eng.latency([s])[0]) == refs[q][1]
        assert np.array_equal(eng.get_tour(s), refs[q][0])
        assert int(rounds[q]) == refs[q][2]


@pytest.mark.parametrize("n,span", [(700, 30000), (2500, 20000)])
def test_rows_int64_path(eng, n, span):
    """Coordinates spread over [0, span]^2 push the |delta| bound past 2^31:
    the tours take the int64 path; results stay identical to the oracle."""
    xs, ys = O.instance_uniform(n, 77, span)
    W = O.weights(xs, ys, 1)
    inst = eng.instance(xs, ys, 1)
    s = eng.slot(inst)
    t = O.perm_tour(78, n)
    eng.set_tour(s, t)
    assert int(eng.latency([s])[0]) == O.latency(W, t)
    for rowmap in (0, 1):
        eng.eval([s], 0x1F, rowmap)
        for nb in range(5):
            _check_rows(eng, s, W, t, nb, rowmap)


@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("n,span", [(5, 2000), (9, 2000), (300, 2000), (1000, 2000),
                                    (2000, 2000), (700, 30000)])
def test_tile_shapes_and_row_buffers(eng, tile, n, span):
    """Every sweep tile shape (column-split rows folded with atomicMin into
    double-buffered row arrays) gives the oracle's rows, across repeated
    evaluations and alternating row maps; span 30000 takes the int64 path."""
    xs, ys = O.instance_clustered(n, 300 + n) if span == 2000 else \
        O.instance_uniform(n, 300 + n, span)
    W = O.weights(xs, ys, 1)
    inst = eng.instance(xs, ys, 1)
    s = eng.slot(inst)
    eng.set_tile(tile)
    try:
        for rep, (rowmap, tseed) in enumerate([(0, 1), (0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (0, 3),
                                               (1, 4), (1, 4), (1, 5)]):
            t = O.perm_tour(tseed + n, n)
            eng.set_tour(s, t)
            eng.eval([s], 0x1F, rowmap)
            for nb in range(5):
                _check_rows(eng, s, W, t, nb, rowmap)
            # a neighborhood subset leaves the others' buffers alone
            eng.eval([s], 0b00101, rowmap)
            for nb in (0, 2):
                _check_rows(eng, s, W, t, nb, rowmap)
            eng.eval([s], 0b11100, rowmap)
            for nb in (2, 3, 4):
                _check_rows(eng, s, W, t, nb, rowmap)
    finally:
        eng.set_tile(-1)


def test_sqrt_exhaustive_vs_library(eng):
    """The kernels' square root equals the library's correctly rounded sqrtf
    for float(s) of every s in [0, 2^31)."""
    assert eng.sqrt_selfcheck() == 0


def test_instance_beyond_64bit_keys_is_refused(eng):
    """n = 14000 over the full coordinate range: N (N + 32) maxd >= 2^43."""
    rng = np.random.RandomState(3)
    xs = rng.randint(0, 32768, 14000).astype(np.int32)
    ys = rng.randint(0, 32768, 14000).astype(np.int32)
    xs[0], ys[0], xs[1], ys[1] = 0, 0, 32767, 32767
    with pytest.raises(wm.MLPError):
        eng.instance(xs, ys, 1)
    inst = eng.instance(xs[:12000], ys[:12000], 1)  # below the bound
    eng.destroy_instance(inst)
