"""Or-opt-k beyond k = 3 (neighborhood ids 5..14, evaluated one wave per row)
against the oracle: per-row bests in both row maps and both integer paths,
merged steps, and DVND over ids 0..9."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

wm = pytest.importorskip("wamca_mi")


@pytest.fixture(scope="module")
def eng():
    e = wm.Engine(0)
    e.set_neighborhoods(15)
    yield e
    e.dvnd_neighborhoods(5)
    e.set_neighborhoods(5)
    e.close()


@pytest.mark.parametrize("rowmap", [0, 1])
@pytest.mark.parametrize("n,span", [(60, 1000), (400, 1000), (700, 30000)])
def test_rows_and_steps(eng, n, span, rowmap):
    xs, ys = O.instance_uniform(n, 90 + n, span)
    W = O.weights(xs, ys, 1)
    inst = eng.instance(xs, ys, 1)
    s = eng.slot(inst)
    try:
        t = O.perm_tour(n + 3, n)
        eng.set_tour(s, t)
        eng.eval([s], 0x7FFF, rowmap)  # tiled ids 0..4 and per-row ids 5..14 in one call
        for nb in range(15):
            if O.rows(nb, rowmap, n + 1) <= 0:
                continue
            ref = O.eval_rows(nb, rowmap, W, t)
            got = eng.rows(s, nb)
            assert np.array_equal(got["cost"], ref["cost"]), nb
            assert np.array_equal(got["i"], ref["i"]) and np.array_equal(got["j"], ref["j"]), nb
        for nb in (5, 9, 14):
            if O.rows(nb, rowmap, n + 1) <= 0:
                continue
            eng.set_tour(s, t)
            eng.eval([s], 1 << nb, rowmap)
            eng.merge_apply([s], nb)
            t2, cnt, _ = O.step(nb, rowmap, W, t)
            assert np.array_equal(eng.get_tour(s), t2), nb
            assert len(eng.applied(s, nb)) == cnt
    finally:
        eng.destroy_slot(s)
        eng.destroy_instance(inst)


@pytest.mark.parametrize("rowmap", [0, 1])
def test_dvnd_ten_neighborhoods(eng, rowmap):
    eng.dvnd_neighborhoods(10)
    insts, slots, refs = [], [], []
    try:
        for q, n in enumerate([70, 150, 260]):
            xs, ys = O.instance_clustered(n, 300 + q) if q % 2 else O.instance_uniform(n, 300 + q)
            W = O.weights(xs, ys, 1)
            t = O.random_tour(O.MT(300 + q), W, 0.5)
            inst = eng.instance(xs, ys, 1)
            s = eng.slot(inst)
            eng.set_tour(s, t)
            insts.append(inst)
            slots.append(s)
            refs.append(O.dvnd(W, t, rowmap, 0, 10))
        rounds = eng.dvnd(slots, rowmap)
        for q, s in enumerate(slots):
            tref, Lref, rref, _ = refs[q]
            assert int(eng.latency([s])[0]) == Lref and int(rounds[q]) == rref
            assert np.array_equal(eng.get_tour(s), tref)
    finally:
        eng.dvnd_neighborhoods(5)
        for s in slots:
            eng.destroy_slot(s)
        for i in insts:
            eng.destroy_instance(i)
