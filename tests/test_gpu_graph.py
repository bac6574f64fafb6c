"""HIP-graph capture of batched evaluations and merges (mlp_graph_*): every
replay equals the eager result, with eager work (new tours, other launches,
synchronisations) between replays; synchronising calls are refused while a
capture is open."""
import contextlib

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

wm = pytest.importorskip("wamca_mi")


@pytest.fixture(scope="module")
def eng():
    e = wm.Engine(0)
    yield e
    e.close()


@contextlib.contextmanager
def capture(eng, out):
    """Record into a graph; the capture is closed even when a call fails."""
    eng.graph_begin()
    try:
        yield
    except Exception:
        try:
            eng.graph_end()
        except wm.MLPError:
            pass
        raise
    out.append(eng.graph_end())


@pytest.mark.parametrize("persistent", [False, True])
def test_eval_graph_replays(eng, persistent):
    n = 200
    xs, ys = O.instance_clustered(n, 77)
    W = O.weights(xs, ys, 1)
    inst = eng.instance(xs, ys, 1)
    S = eng.slots(inst, 40)
    try:
        eng.set_tours(S, [O.perm_tour(q + 1, n) for q in range(40)])
        eng.sync()
        gs = []
        with capture(eng, gs):
            eng.eval(S, 0x1F, 0, persistent=persistent)
        g = gs[0]
        for rep in range(4):
            tours = [O.perm_tour(100 * rep + q + 7, n) for q in range(40)]
            eng.set_tours(S, tours)
            eng.graph_launch(g)
            eng.eval(S[:3], 0x04, 0)  # eager work on some of the slots in between
            eng.sync()
            for q in (0, 5, 39):
                for nb in (1, 3) if q < 3 else range(5):
                    ref = O.eval_rows(nb, 0, W, tours[q])
                    assert np.array_equal(eng.rows(S[q], nb)["cost"], ref["cost"]), (rep, q, nb)
        eng.graph_destroy(g)
    finally:
        for s in S:
            eng.destroy_slot(s)
        eng.destroy_instance(inst)


def test_step_graph_replays(eng):
    """A graph of {eval swap, merge swap, eval or-opt-1, merge} replayed three
    times equals three eager applications of the oracle's steps."""
    n = 150
    xs, ys = O.instance_uniform(n, 21)
    W = O.weights(xs, ys, 1)
    inst = eng.instance(xs, ys, 1)
    S = eng.slots(inst, 3)
    try:
        tours = [O.random_tour(O.MT(60 + q), W, 0.5) for q in range(3)]
        eng.set_tours(S, tours)
        eng.sync()
        gs = []
        with capture(eng, gs):
            for nb in (0, 2):
                eng.eval(S, 1 << nb, 0)
                eng.merge_apply(S, nb)
        g = gs[0]
        refs = [t.copy() for t in tours]
        for rep in range(3):
            eng.graph_launch(g)
            eng.sync()
            for q in range(3):
                for nb in (0, 2):
                    refs[q] = O.step(nb, 0, W, refs[q])[0]
                assert np.array_equal(eng.get_tour(S[q]), refs[q]), (rep, q)
        eng.graph_destroy(g)
    finally:
        for s in S:
            eng.destroy_slot(s)
        eng.destroy_instance(inst)


def test_capture_refuses_sync_calls(eng):
    xs, ys = O.instance_uniform(50, 3)
    inst = eng.instance(xs, ys, 1)
    s = eng.slot(inst)
    try:
        eng.set_tour(s, O.perm_tour(1, 50))
        gs = []
        with capture(eng, gs):
            eng.eval([s], 1, 0)
            with pytest.raises(wm.MLPError):
                eng.rows(s, 0)
        g = gs[0]
        eng.graph_launch(g)
        eng.sync()
        W = O.weights(xs, ys, 1)
        assert np.array_equal(eng.rows(s, 0)["cost"], O.eval_rows(0, 0, W, O.perm_tour(1, 50))["cost"])
        eng.graph_destroy(g)
    finally:
        eng.destroy_slot(s)
        eng.destroy_instance(inst)
