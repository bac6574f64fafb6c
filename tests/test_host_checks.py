"""Host-side input checks of the C ABI (no GPU): mlp_check_job applies the
checks instance creation and tour upload apply, so a serving loop can refuse a
job before queueing it."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

wm = pytest.importorskip("wamca_mi")
from wamca_mi._lib import lib  # noqa: E402


def _check(xs, ys, tour, rnd=1):
    xs = np.ascontiguousarray(xs, np.int32)
    ys = np.ascontiguousarray(ys, np.int32)
    return lib().mlp_check_job(None, xs, ys, len(xs), rnd,
                               np.ascontiguousarray(tour, np.int32))


def test_check_job_statuses():
    xs, ys = O.instance_uniform(100, 1)
    t = O.perm_tour(1, 100)
    assert _check(xs, ys, t) == 0
    # negative coordinates with a span inside 32767 are accepted (translated)
    assert _check(xs - 5000, ys - 20000, t) == 0
    wide = xs.copy()
    wide[7] = int(xs.min()) + 32768
    assert _check(wide, ys, t) == -4
    assert _check(xs, ys, t, rnd=2) == -1
    bad = t.copy()
    bad[10] = bad[11]
    assert _check(xs, ys, bad) == -1
    bad = t.copy()
    bad[-1] = 5
    assert _check(xs, ys, bad) == -1
    bx, by = O.instance_uniform(14000, 3, 32767)
    assert _check(bx, by, O.perm_tour(3, 14000)) == -4
    # the same size on a small square fits the 64-bit keys
    sx, sy = O.instance_uniform(14000, 3, 1000)
    assert _check(sx, sy, O.perm_tour(3, 14000)) == 0
    assert lib().mlp_check_job(None, xs[:1].copy(), ys[:1].copy(), 1, 1, t[:2].copy()) == -1
