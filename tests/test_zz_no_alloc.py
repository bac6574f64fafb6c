"""mlp_eval_static (the entry point that allocates nothing) against mlp_eval +
mlp_get_rows on the same instance and tour: two calls in a row, the second on
a smaller tour, both row maps; every row's cost, i and j must be equal."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

wm = pytest.importorskip("wamca_mi")

STATIC_MAX_N = 8192


class DevArray:
    """Device array of the engine's context (mlp_dev_alloc), for test set-up."""

    def __init__(self, eng, dtype, count):
        self.eng, self.dtype, self.count = eng, np.dtype(dtype), count
        p = C.c_void_p()
        eng._ck(eng._L.mlp_dev_alloc(eng.h, self.dtype.itemsize * count, C.byref(p)))
        self.p = p

    def put(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.size <= self.count
        self.eng._ck(self.eng._L.mlp_copy_to_device(self.eng.h, self.p, a.ctypes.data_as(C.c_void_p),
                                                    a.nbytes))

    def get(self, count):
        out = np.empty(count, self.dtype)
        self.eng._ck(self.eng._L.mlp_copy_to_host(self.eng.h, out.ctypes.data_as(C.c_void_p), self.p,
                                                  out.nbytes))
        return out

    def free(self):
        self.eng._ck(self.eng._L.mlp_dev_free(self.eng.h, self.p))


@pytest.fixture(scope="module")
def eng():
    e = wm.Engine(0)
    yield e
    e.close()


def _static_rows(eng, inst, tour, mask, rowmap, bufs):
    N = len(tour)
    total = sum(wm.rows_of(nb, rowmap, N) for nb in range(5) if (mask >> nb) & 1)
    t_dev, c_dev, i_dev, j_dev = bufs
    t_dev.put(tour)
    eng._ck(eng._L.mlp_eval_static(eng.h, inst, t_dev.p, mask, rowmap, c_dev.p, i_dev.p, j_dev.p))
    return c_dev.get(total), i_dev.get(total), j_dev.get(total)


def test_static_eval_matches_eval(eng):
    cap_rows = 5 * STATIC_MAX_N
    bufs = (DevArray(eng, np.int32, STATIC_MAX_N), DevArray(eng, np.int64, cap_rows),
            DevArray(eng, np.int32, cap_rows), DevArray(eng, np.int32, cap_rows))
    try:
        cases = [(1000, 1, 0x1F, 0, 11), (301, 0, 0x1F, 1, 12), (150, 1, 0x0B, 0, 13),
                 (60, 1, 0x14, 1, 14), (700, 1, 0x1F, 1, 15)]
        for n, rnd, mask, rowmap, seed in cases:
            xs, ys = O.instance_clustered(n, seed) if seed % 2 else O.instance_uniform(n, seed)
            tour = O.perm_tour(seed, n)
            inst = eng.instance(xs, ys, rnd)
            s = eng.slot(inst)
            eng.set_tour(s, tour)
            eng.eval([s], mask, rowmap)
            cost, ii, jj = _static_rows(eng, inst, tour, mask, rowmap, bufs)
            off = 0
            for nb in range(5):
                if not (mask >> nb) & 1:
                    continue
                ref = eng.rows(s, nb)
                R = len(ref["cost"])
                has = ref["cost"] != np.iinfo(np.int64).max
                assert np.array_equal(cost[off:off + R], ref["cost"]), (n, nb)
                assert np.array_equal(ii[off:off + R][has], ref["i"][has]), (n, nb)
                assert np.array_equal(jj[off:off + R][has], ref["j"][has]), (n, nb)
                off += R
            eng.destroy_slot(s)
            eng.destroy_instance(inst)
        # refused: a tour above the static maximum, an or-opt-k id beyond 3
        xs, ys = O.instance_uniform(STATIC_MAX_N, 3)
        big = eng.instance(xs, ys, 1)
        with pytest.raises(wm.MLPError) as e:
            eng._ck(eng._L.mlp_eval_static(eng.h, big, bufs[0].p, 0x1F, 0, bufs[1].p, bufs[2].p,
                                           bufs[3].p))
        assert e.value.status == -4
        eng.destroy_instance(big)
        xs, ys = O.instance_uniform(50, 3)
        small = eng.instance(xs, ys, 1)
        with pytest.raises(wm.MLPError):
            eng._ck(eng._L.mlp_eval_static(eng.h, small, bufs[0].p, 0x20, 0, bufs[1].p,
                                           bufs[2].p, bufs[3].p))
        eng.destroy_instance(small)
    finally:
        for b in bufs:
            b.free()
