"""TOTAL mode on the GPU vs the oracle: per-(row, thread) entries for any
block size, the greedy merge over all entries, the applied tour, and the
logged TOTAL records of the reference's runs (best entry, improving-entry
count, GPU-GPU merged sum and count) recomputed on the device."""
import numpy as np
import pytest

from helpers import load_tsplib_coords, results_logs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

wm = pytest.importorskip("wamca_mi")
INF = np.iinfo(np.int64).max


@pytest.fixture(scope="module")
def eng():
    e = wm.Engine(0)
    e.set_neighborhoods(15)
    yield e
    e.close()


def _check_entries(got, ref, tag):
    assert len(got["cost"]) == len(ref["cost"]), tag
    empty = ref["cost"] == INF
    bad = np.nonzero((got["cost"] != ref["cost"]) |
                     (((got["i"] != ref["i"]) | (got["j"] != ref["j"])) & ~empty))[0]
    assert bad.size == 0, (tag, int(bad[0]), int(got["cost"][bad[0]]), int(ref["cost"][bad[0]]))


def _check_merge(eng, s, W, t, nb, ref):
    """merge_total on slot s (entries already evaluated) against the oracle."""
    kept, ssum = O.merge(nb, ref["cost"], ref["i"], ref["j"])
    L0 = int(eng.latency([s])[0])
    eng.merge_total([s])
    got = eng.applied(s, nb)
    assert [g[4] for g in got] == kept.tolist()
    assert [(g[0], g[2], g[3]) for g in got] == [
        (int(ref["cost"][q]), int(ref["i"][q]), int(ref["j"][q])) for q in kept]
    t2 = t.copy()
    for q in kept:
        t2 = O.apply(nb, t2, int(ref["i"][q]), int(ref["j"][q]))
    assert np.array_equal(eng.get_tour(s), t2)
    assert int(eng.latency([s])[0]) == L0 + ssum == O.latency(W, t2)
    return len(kept), ssum


def _block(k, N):
    p = k["total"]
    return p["block"] if "block" in p else p["move_elems"] // O.rows(k["nb"], 0, N)


def test_logged_total_records(eng):
    """Every logged TOTAL record, recomputed on the device for the logged block."""
    done = 0
    for log in results_logs():
        xs, ys = load_tsplib_coords(log["instance"])
        W = O.weights(xs, ys, 0)
        inst = eng.instance(xs, ys, 0)
        s = eng.slot(inst)
        for sol in log["solutions"]:
            t = np.array(sol["tour"], np.int32)
            for name, k in sol["kernels"].items():
                p = k.get("total", {})
                if "best" not in p:
                    continue
                eng.set_tour(s, t)
                B = _block(k, len(t))
                eng.eval_total([s], k["nb"], B)
                r = eng.total(s)
                assert len(r["cost"]) == p["move_elems"]
                b = int(np.argsort(r["cost"], kind="stable")[0])
                assert [int(r["i"][b]), int(r["j"][b]), int(r["cost"][b])] == p["best"]
                v = int((r["cost"] < 0).sum())
                assert v == p["graph"][0] or (p["graph"][0] == 120000 and v > 120000)
                if "merge_gpu" in p:
                    L0 = int(eng.latency([s])[0])
                    eng.merge_total([s])
                    got = eng.applied(s, k["nb"])
                    ssum = sum(g[0] for g in got)
                    assert [ssum, len(got)] == p["merge_gpu"], (log["log"], sol["index"], name)
                    assert int(eng.latency([s])[0]) == L0 + ssum
                done += 1
        eng.destroy_slot(s)
        eng.destroy_instance(inst)
    assert done == 66


@pytest.mark.parametrize("n,rnd,seed", [(40, 0, 1), (101, 1, 2), (300, 1, 3)])
@pytest.mark.parametrize("block", [1, 7, 64, 250, 256, 1000])
def test_entries_and_merge_vs_oracle(eng, n, rnd, seed, block):
    xs, ys = O.instance_uniform(n, seed)
    W = O.weights(xs, ys, rnd)
    inst = eng.instance(xs, ys, rnd)
    s = eng.slot(inst)
    for nb in range(5):
        t = O.perm_tour(seed * 10 + nb, n)
        eng.set_tour(s, t)
        eng.eval_total([s], nb, block)
        ref = O.eval_total(nb, W, t, block)
        _check_entries(eng.total(s), ref, (n, block, nb))
        _check_merge(eng, s, W, t, nb, ref)
    eng.destroy_slot(s)
    eng.destroy_instance(inst)


@pytest.mark.parametrize("n,seed", [(30, 4), (120, 5)])
def test_oropt_k_beyond_3(eng, n, seed):
    xs, ys = O.instance_uniform(n, seed)
    W = O.weights(xs, ys, 1)
    inst = eng.instance(xs, ys, 1)
    s = eng.slot(inst)
    for nb in range(5, 15):
        t = O.perm_tour(seed + nb, n)
        eng.set_tour(s, t)
        eng.eval_total([s], nb, 16)
        ref = O.eval_total(nb, W, t, 16)
        _check_entries(eng.total(s), ref, nb)
        if len(ref["cost"]):
            _check_merge(eng, s, W, t, nb, ref)
    eng.destroy_slot(s)
    eng.destroy_instance(inst)


def test_every_move_its_own_entry(eng):
    """block >= the row length: the entries are all moves of the row map;
    the merge is then the greedy over every improving move."""
    n = 200
    xs, ys = O.instance_uniform(n, 8)
    W = O.weights(xs, ys, 1)
    inst = eng.instance(xs, ys, 1)
    s = eng.slot(inst)
    t = O.random_tour(O.MT(8), W, 0.5)
    for nb in range(5):
        eng.set_tour(s, t)
        eng.eval_total([s], nb, 256)
        ref = O.eval_total(nb, W, t, 256)
        _check_entries(eng.total(s), ref, nb)
        m = O.enumerate_moves(nb, 0, W, t)
        assert int((ref["cost"] != INF).sum()) == len(m["ads"])
        _check_merge(eng, s, W, t, nb, ref)
    eng.destroy_slot(s)
    eng.destroy_instance(inst)


def test_batch_global_arrays_and_tiny(eng):
    """Several slots in one call, the device-memory variant of the merge, and
    tours too short to have moves."""
    xs, ys = O.instance_uniform(500, 11)
    W = O.weights(xs, ys, 1)
    inst = eng.instance(xs, ys, 1)
    ss = eng.slots(inst, 3)
    tours = [O.perm_tour(50 + q, 500) for q in range(3)]
    for glob in (False, True):
        eng.merge_force_global(glob)
        try:
            eng.set_tours(ss, tours)
            eng.eval_total(ss, 1, 512)
            refs = [O.eval_total(1, W, t, 512) for t in tours]
            for s, t, ref in zip(ss, tours, refs):
                _check_entries(eng.total(s), ref, glob)
            kept = [O.merge(1, r["cost"], r["i"], r["j"]) for r in refs]
            eng.merge_total(ss)
            for s, t, ref, (kp, ssum) in zip(ss, tours, refs, kept):
                got = eng.applied(s, 1)
                assert [g[4] for g in got] == kp.tolist()
                assert int(eng.latency([s])[0]) == O.latency(W, t) + ssum
        finally:
            eng.merge_force_global(False)
    for n in (5, 6, 9):
        xs, ys = O.instance_uniform(n, n, 50)
        W = O.weights(xs, ys, 0)
        inst = eng.instance(xs, ys, 0)
        s = eng.slot(inst)
        for nb in range(15):
            t = O.perm_tour(n + nb, n)
            eng.set_tour(s, t)
            eng.eval_total([s], nb, 3)
            ref = O.eval_total(nb, W, t, 3)
            _check_entries(eng.total(s), ref, (n, nb))
            _check_merge(eng, s, W, t, nb, ref)


def test_errors(eng):
    xs, ys = O.instance_uniform(20, 1)
    inst = eng.instance(xs, ys, 0)
    s = eng.slot(inst)
    eng.set_tour(s, O.perm_tour(1, 20))
    with pytest.raises(wm.MLPError):
        eng.merge_total([s])  # nothing evaluated
    with pytest.raises(wm.MLPError):
        eng.eval_total([s], 0, 0)
    with pytest.raises(wm.MLPError):
        eng.eval_total([s], 0, 16385)
    eng.eval_total([s], 0, 8)
    eng.merge_total([s])
    with pytest.raises(wm.MLPError):
        eng.merge_total([s])  # already merged
    eng.set_tour(s, O.perm_tour(2, 20))
    with pytest.raises(wm.MLPError):
        eng.merge_total([s])  # the entries belonged to the previous tour
