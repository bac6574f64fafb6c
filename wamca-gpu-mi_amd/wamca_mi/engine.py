"""Python view of the C ABI (include/mlp_gpu.h): one Engine per HIP context.

Thin glue for tests and the benchmark: every call goes straight to
``libwamca_mi.so``; errors raise :class:`MLPError` with the library's message.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import I32P, MLPError, MoveRec, lib  # noqa: F401

NB_NAMES = ("swap", "2opt", "oropt1", "oropt2", "oropt3")
ROWMAP_ADS, ROWMAP_ODC = 0, 1


MAX_N = 32768  # MLP_MAX_N: largest tour size N (positions)
N_TILES = 5    # sweep tile shapes


def rows_of(nb: int, rowmap: int, N: int) -> int:
    """Rows of a neighborhood's launch grid (mlkswap.cu:991,1003; mlk2opt.cu:661;
    mlkoropt.cu:819)."""
    if nb == 0:
        return N - 4 if rowmap == 0 else (N + 1) // 2 - 1
    if nb == 1:
        return (N + 1) // 2
    return N - (nb - 1) - 2


class Engine:
    def __init__(self, device: int = 0, tuned: bool = True):
        """Context on one GPU.  tuned: load the committed tile-shape table of
        this architecture (wamca_mi/tune_<arch>.json) when there is one."""
        self._L = lib()
        h = C.c_void_p()
        st = self._L.mlp_ctx_create(device, C.byref(h))
        if st != 0:
            raise MLPError(st, self._L.mlp_strerror(st).decode())
        self.h = h
        self.N = {}  # slot -> N
        if tuned:
            from .tune import load_table
            t = load_table(self.device_info()["arch"])
            if t is not None:
                self.set_tune_table(t)

    def close(self):
        if self.h:
            self._L.mlp_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, st):
        if st != 0:
            raise MLPError(st, self._L.mlp_ctx_error(self.h).decode())

    # -- device -------------------------------------------------------------
    def device_info(self):
        cu = C.c_int32()
        hbm = C.c_int64()
        name = C.create_string_buffer(256)
        self._ck(self._L.mlp_device_info(self.h, C.byref(cu), C.byref(hbm), name, 256))
        return {"cus": cu.value, "hbm_bytes": hbm.value, "arch": name.value.decode()}

    def sync(self):
        self._ck(self._L.mlp_sync(self.h))

    # -- instances / slots --------------------------------------------------
    def instance(self, xs, ys, rnd=0) -> int:
        xs = np.ascontiguousarray(xs, np.int32)
        ys = np.ascontiguousarray(ys, np.int32)
        out = C.c_int32()
        self._ck(self._L.mlp_instance_create(self.h, xs, ys, len(xs), rnd, C.byref(out)))
        self._n_of = getattr(self, "_n_of", {})
        self._n_of[out.value] = len(xs)
        return out.value

    def destroy_instance(self, inst):
        self._ck(self._L.mlp_instance_destroy(self.h, inst))

    def weights(self, inst) -> np.ndarray:
        n = self._n_of[inst]
        w = np.empty(n * n, np.int32)
        self._ck(self._L.mlp_instance_weights(self.h, inst, w))
        return w.reshape(n, n)

    def slot(self, inst) -> int:
        out = C.c_int32()
        self._ck(self._L.mlp_slot_create(self.h, inst, C.byref(out)))
        self.N[out.value] = self._n_of[inst] + 1
        return out.value

    def slots(self, inst, count):
        return [self.slot(inst) for _ in range(count)]

    def destroy_slot(self, s):
        self._ck(self._L.mlp_slot_destroy(self.h, s))
        self.N.pop(s, None)

    @staticmethod
    def footprint(n) -> int:
        return int(lib().mlp_slot_footprint(n))

    # -- tours --------------------------------------------------------------
    def set_tours(self, slots, tours):
        slots = np.ascontiguousarray(slots, np.int32)
        flat = np.ascontiguousarray(np.concatenate([np.asarray(t, np.int32) for t in tours]))
        self._ck(self._L.mlp_set_tours(self.h, slots, len(slots), flat))

    def set_tour(self, s, tour):
        self.set_tours([s], [tour])

    def get_tours(self, slots):
        slots = np.ascontiguousarray(slots, np.int32)
        tot = sum(self.N[int(s)] for s in slots)
        out = np.empty(tot, np.int32)
        self._ck(self._L.mlp_get_tours(self.h, slots, len(slots), out))
        res, off = [], 0
        for s in slots:
            res.append(out[off:off + self.N[int(s)]].copy())
            off += self.N[int(s)]
        return res

    def get_tour(self, s):
        return self.get_tours([s])[0]

    def latency(self, slots):
        slots = np.ascontiguousarray(slots, np.int32)
        out = np.empty(len(slots), np.int64)
        self._ck(self._L.mlp_get_latency(self.h, slots, len(slots), out))
        return out

    # -- neighborhoods ------------------------------------------------------
    def eval(self, slots, nb_mask=0x1F, rowmap=ROWMAP_ADS, persistent=False):
        slots = np.ascontiguousarray(slots, np.int32)
        self._ck(self._L.mlp_eval(self.h, slots, len(slots), nb_mask, rowmap, int(persistent)))

    def rows(self, s, nb):
        cap = MAX_N
        cost = np.empty(cap, np.int64)
        i = np.empty(cap, np.int32)
        j = np.empty(cap, np.int32)
        R = C.c_int32()
        self._ck(self._L.mlp_get_rows(self.h, s, nb, cost, i, j, C.byref(R)))
        r = R.value
        return {"cost": cost[:r].copy(), "i": i[:r].copy(), "j": j[:r].copy()}

    def moves64(self, s, nb):
        out = np.empty(MAX_N, np.uint64)
        R = C.c_int32()
        self._ck(self._L.mlp_get_moves64(self.h, s, nb, out, C.byref(R)))
        return out[:R.value].copy()

    def merge_apply(self, slots, nb):
        slots = np.ascontiguousarray(slots, np.int32)
        self._ck(self._L.mlp_merge_apply(self.h, slots, len(slots), nb))

    # -- TOTAL mode (per-(row, thread) entries) ---------------------------------
    def eval_total(self, slots, nb, block):
        """Entries row * block + t: best move of nb among the row-local
        indices t, t + block, ... (the reference's TOTAL kernels)."""
        slots = np.ascontiguousarray(slots, np.int32)
        self._ck(self._L.mlp_eval_total(self.h, slots, len(slots), nb, block))

    def total(self, s):
        M = C.c_int64()
        self._ck(self._L.mlp_get_total(self.h, s, None, None, None, 0, C.byref(M)))
        m = M.value
        cost = np.empty(m, np.int64)
        i = np.empty(m, np.int32)
        j = np.empty(m, np.int32)
        if m:
            self._ck(self._L.mlp_get_total(self.h, s, cost.ctypes.data, i.ctypes.data,
                                           j.ctypes.data, m, C.byref(M)))
        return {"cost": cost, "i": i, "j": j}

    def merge_total(self, slots):
        """Greedy merge over all TOTAL entries + apply (kept moves: applied())."""
        slots = np.ascontiguousarray(slots, np.int32)
        self._ck(self._L.mlp_merge_total(self.h, slots, len(slots)))

    def applied(self, s, nb):
        cap = MAX_N
        buf = (MoveRec * cap)()
        cnt = C.c_int32()
        self._ck(self._L.mlp_get_applied(self.h, s, nb, buf, cap, C.byref(cnt)))
        return [(buf[q].cost, buf[q].id, buf[q].i, buf[q].j, buf[q].row) for q in range(cnt.value)]

    # -- HIP graphs -----------------------------------------------------------
    def graph_begin(self):
        """Start recording evals / merges / DVND rounds into a graph."""
        self._ck(self._L.mlp_graph_begin(self.h))

    def graph_end(self) -> int:
        g = C.c_int32()
        self._ck(self._L.mlp_graph_end(self.h, C.byref(g)))
        return g.value

    def graph_launch(self, g: int):
        self._ck(self._L.mlp_graph_launch(self.h, g))

    def graph_destroy(self, g: int):
        self._ck(self._L.mlp_graph_destroy(self.h, g))

    def set_neighborhoods(self, count: int):
        """Slots created from now on hold neighborhood ids 0..count-1 (ids
        5..14 are or-opt-4..13)."""
        self._ck(self._L.mlp_set_neighborhoods(self.h, count))

    def dvnd_neighborhoods(self, count: int):
        """A DVND round runs neighborhood ids 0..count-1 (default 5)."""
        self._ck(self._L.mlp_dvnd_neighborhoods(self.h, count))

    def slot_bytes(self, n: int) -> int:
        b = C.c_int64()
        self._ck(self._L.mlp_slot_bytes(self.h, n, C.byref(b)))
        return b.value

    def merge_force_global(self, on: bool):
        """Merge working arrays in device memory for every tour size."""
        self._ck(self._L.mlp_merge_force_global(self.h, int(on)))

    def dvnd(self, slots, rowmap=ROWMAP_ADS, max_rounds=0):
        slots = np.ascontiguousarray(slots, np.int32)
        rounds = np.zeros(len(slots), np.int32)
        self._ck(self._L.mlp_dvnd(self.h, slots, len(slots), rowmap, max_rounds, rounds))
        return rounds

    def dvnd_config(self, chunk=4, use_graph=True, persistent=False):
        """Rounds per convergence check, HIP-graph replay of a chunk, persistent grids."""
        self._ck(self._L.mlp_dvnd_config(self.h, chunk, int(use_graph), int(persistent)))

    def dvnd_fused(self, on=True):
        """Each DVND round as one persistent launch (a dependency-ordered
        queue of every slot's evaluation tiles and merges)."""
        self._ck(self._L.mlp_dvnd_fused(self.h, int(on)))

    def dvnd_groups(self, groups=2):
        """mlp_dvnd over two slot groups on two streams (the largest slots,
        and the rest), each checked for convergence when its own chunk of
        rounds ends, or one group (1)."""
        self._ck(self._L.mlp_dvnd_groups(self.h, int(groups)))

    def dvnd_start(self, slots):
        slots = np.ascontiguousarray(slots, np.int32)
        self._ck(self._L.mlp_dvnd_start(self.h, slots, len(slots)))

    def dvnd_rounds(self, slots, rounds, rowmap=ROWMAP_ADS, max_rounds=0):
        """Enqueue `rounds` DVND rounds over the slots (asynchronous)."""
        slots = np.ascontiguousarray(slots, np.int32)
        self._ck(self._L.mlp_dvnd_rounds(self.h, slots, len(slots), rowmap, rounds, max_rounds))

    def dvnd_status(self, slots):
        """(active flags, rounds run) per slot; synchronises the stream."""
        slots = np.ascontiguousarray(slots, np.int32)
        act = np.zeros(len(slots), np.int32)
        rnd = np.zeros(len(slots), np.int32)
        self._ck(self._L.mlp_dvnd_status(self.h, slots, len(slots), act, rnd))
        return act, rnd

    def ads(self, s):
        N = self.N[s]
        nbytes = 128 + 4 * N * (N + 3)
        buf = np.empty(nbytes // 4, np.uint32)
        self._ck(self._L.mlp_ads_materialize(self.h, s, buf, nbytes))
        return buf

    def dist_table(self, s0, count, rnd):
        out = np.empty(count, np.int32)
        self._ck(self._L.mlp_dist_table(self.h, s0, count, rnd, out))
        return out

    def dist_list(self, s, rnd):
        s = np.ascontiguousarray(s, np.int32)
        out = np.empty(len(s), np.int32)
        self._ck(self._L.mlp_dist_list(self.h, s, len(s), rnd, out))
        return out

    def sqrt_selfcheck(self) -> int:
        out = C.c_int64()
        self._ck(self._L.mlp_sqrt_selfcheck(self.h, C.byref(out)))
        return out.value

    def set_tile(self, tile: int):
        """Sweep tile shape: -1 automatic, 0 = 4x256, 1 = 8x512, 2 = 16x1024,
        3 = 8x256, 4 = 4x512."""
        self._ck(self._L.mlp_set_tile(self.h, tile))

    def autotune(self, slots, nb_mask=0x1F, reps=3):
        """Time every sweep tile shape on this launch; record the best in the
        context's table.  Returns (best shape, [us per launch per shape])."""
        slots = np.ascontiguousarray(slots, np.int32)
        best = C.c_int32()
        us = (C.c_float * N_TILES)()
        self._ck(self._L.mlp_autotune(self.h, slots, len(slots), nb_mask, reps, C.byref(best), us))
        return best.value, list(us)

    def tune_table(self) -> np.ndarray:
        t = np.empty((32, 48), np.int8)
        self._ck(self._L.mlp_tune_get(self.h, t))
        return t

    def set_tune_table(self, t):
        t = np.ascontiguousarray(t, np.int8).reshape(32, 48)
        self._ck(self._L.mlp_tune_set(self.h, t))

    def timing_begin(self, capacity=8192):
        self._ck(self._L.mlp_timing_begin(self.h, capacity))

    def timing_end(self):
        """(summed eval kernel ms, launches) since timing_begin."""
        tot = C.c_double()
        cnt = C.c_int32()
        self._ck(self._L.mlp_timing_end(self.h, C.byref(tot), C.byref(cnt)))
        return tot.value, cnt.value

    def region_begin(self):
        """Open a timed region on the context stream (HIP event)."""
        self._ck(self._L.mlp_region_begin(self.h))

    def region_end(self) -> float:
        """Close the timed region; synchronises; elapsed milliseconds."""
        ms = C.c_float()
        self._ck(self._L.mlp_region_end(self.h, C.byref(ms)))
        return ms.value

    def eval_each(self, slots, nb_mask=0x1F, rowmap=ROWMAP_ADS):
        """One eval launch per slot, enqueued from a C loop."""
        slots = np.ascontiguousarray(slots, np.int32)
        self._ck(self._L.mlp_eval_each(self.h, slots, len(slots), nb_mask, rowmap))

    def copy_probe(self, nbytes, reps=1):
        """Device-to-device copies of nbytes (counter calibration)."""
        self._ck(self._L.mlp_copy_probe(self.h, int(nbytes), int(reps)))

    def launch_probe(self, slot, kind, grid, reps=1):
        """Fixed-cost probe launches (kind 0 empty .. 3 with a tile's loads
        and atomics) with a sweep's launch shape."""
        self._ck(self._L.mlp_launch_probe(self.h, slot, kind, grid, reps))
