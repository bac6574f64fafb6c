"""Large tours: the merge's working arrays in device memory (beyond LDS), the
per-neighborhood int32/int64 choice (swap and or-opt stay on the int32 tiles
while 2-opt takes the int64 path), and tours up to n = 18000 -- each step
compared with the oracle's step on the same tour."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

wm = pytest.importorskip("wamca_mi")


@pytest.fixture(scope="module")
def eng():
    e = wm.Engine(0)
    yield e
    e.merge_force_global(False)
    e.close()


@pytest.mark.parametrize("rowmap", [0, 1])
@pytest.mark.parametrize("n", [52, 300, 1000])
def test_global_merge_equals_oracle(eng, n, rowmap):
    xs, ys = O.instance_clustered(n, 500 + n)
    W = O.weights(xs, ys, 1)
    inst = eng.instance(xs, ys, 1)
    s = eng.slot(inst)
    eng.merge_force_global(True)
    try:
        t = O.random_tour(O.MT(n), W, 0.5)
        for nb in range(5):
            eng.set_tour(s, t)
            eng.eval([s], 1 << nb, rowmap)
            eng.merge_apply([s], nb)
            t2, cnt, tot = O.step(nb, rowmap, W, t)
            assert np.array_equal(eng.get_tour(s), t2), nb
            assert len(eng.applied(s, nb)) == cnt
        eng.set_tour(s, t)
        rounds = eng.dvnd([s], rowmap)
        tref, Lref, rref, _ = O.dvnd(W, t, rowmap)
        assert int(eng.latency([s])[0]) == Lref and int(rounds[0]) == rref
        assert np.array_equal(eng.get_tour(s), tref)
    finally:
        eng.merge_force_global(False)
        eng.destroy_slot(s)
        eng.destroy_instance(inst)


def _step_check(eng, s, W, t, nbs, rowmap=0):
    for nb in nbs:
        eng.set_tour(s, t)
        eng.eval([s], 1 << nb, rowmap)
        got = eng.rows(s, nb)
        ref = O.eval_rows(nb, rowmap, W, t)
        assert np.array_equal(got["cost"], ref["cost"]), nb
        eng.merge_apply([s], nb)
        t2, cnt, _ = O.step(nb, rowmap, W, t)
        assert np.array_equal(eng.get_tour(s), t2), nb


def test_n9000_rows_beyond_lds(eng):
    """n = 9000 uniform: more rows than the LDS merge holds; 2-opt deltas of a
    random permutation exceed int32 (int64 path), swap / or-opt do not."""
    n = 9000
    xs, ys = O.instance_uniform(n, 9)
    W = O.weights(xs, ys, 1)
    inst = eng.instance(xs, ys, 1)
    s = eng.slot(inst)
    try:
        _step_check(eng, s, W, O.perm_tour(9, n), range(5))
    finally:
        eng.destroy_slot(s)
        eng.destroy_instance(inst)


def test_n18000_clustered(eng):
    """n = 18000 clustered (the size of the serving stream's large jobs), a
    random-permutation tour: swap and or-opt-1 steps equal the oracle's; a few
    DVND rounds keep the tour a permutation whose latency the host recomputes
    identically."""
    n = 18000
    xs, ys = O.instance_clustered(n, 18)
    W = O.weights(xs, ys, 1)
    inst = eng.instance(xs, ys, 1)
    s = eng.slot(inst)
    try:
        t = O.perm_tour(18, n)
        _step_check(eng, s, W, t, (0, 2))
        eng.set_tour(s, t)
        eng.dvnd([s], 0, 3)
        tf = eng.get_tour(s)
        assert tf[0] == 0 and tf[-1] == 0 and sorted(tf[1:-1].tolist()) == list(range(1, n))
        assert int(eng.latency([s])[0]) == O.latency(W, tf)
        assert int(eng.latency([s])[0]) < O.latency(W, t)
    finally:
        eng.destroy_slot(s)
        eng.destroy_instance(inst)
