"""Admission by HBM footprint (wloop_amd.serve.Admission), on the CPU."""
import threading
import time

import pytest

import wloop_amd as W
from wloop_amd.serve import Admission, Scheduler, expected_ms, standard_request


class H:
    def __init__(self, name):
        self.name = name
        self.closed = False


def close(h):
    h.closed = True


def test_reuse_of_idle_context():
    a = Admission(100)
    h, create = a.acquire("A", 60, close)
    assert h is None and create
    ha = H("a")
    a.register(ha, "A", 60)
    a.release(ha)
    h, create = a.acquire("A", 60, close)
    assert h is ha and not create and a.used == 60


def test_lru_eviction_of_idle_contexts():
    a = Admission(100)
    for key in ("A", "B"):
        a.acquire(key, 40, close)
        h = H(key)
        a.register(h, key, 40)
        a.release(h)
    first = next(iter(a.idle))
    h, create = a.acquire("C", 50, close)
    assert create and first.closed and a.evicted == 1 and a.used == 90


def test_request_larger_than_budget_is_rejected():
    a = Admission(100)
    with pytest.raises(MemoryError):
        a.acquire("X", 101, close)


def test_waits_for_busy_context():
    a = Admission(100)
    a.acquire("A", 80, close)
    ha = H("a")
    a.register(ha, "A", 80)
    got = {}

    def other():
        got["r"] = a.acquire("B", 50, close)

    t = threading.Thread(target=other)
    t.start()
    time.sleep(0.2)
    assert t.is_alive()                 # blocked: 80 busy + 50 > 100
    a.release(ha)                       # becomes idle -> evicted for B
    t.join(2)
    assert not t.is_alive() and got["r"] == (None, True) and ha.closed


def test_abandon_returns_bytes():
    a = Admission(100)
    a.acquire("A", 70, close)
    a.abandon(70)
    assert a.used == 0


def _req(rid, L, t_arrival, precision=None):
    r = standard_request(rid, L, 3, W.MIXED if precision is None else precision)
    r.t_arrival = t_arrival
    return r


def test_expected_cost_grows_with_size_and_precision():
    costs = [expected_ms(_req(i, L, 0.0)) for i, L in enumerate((12, 16, 24, 32, 48, 64))]
    assert costs == sorted(costs)
    assert 500 < costs[4] < 2000  # 48^3x96, mixed: ~1 s measured
    assert expected_ms(_req(9, 48, 0.0, W.FP64)) > costs[4]


def test_fifo_keeps_arrival_order():
    s = Scheduler("fifo", lambda key: True)
    now = time.perf_counter()
    for i, L in enumerate((64, 12, 32)):
        s.put(_req(i, L, now))
    assert [s.get().rid for _ in range(3)] == [0, 1, 2]


def test_sjf_serves_least_cost_first_and_ages():
    s = Scheduler("sjf", lambda key: True, age_weight=0.05)
    now = time.perf_counter()
    s.put(_req(0, 64, now))
    s.put(_req(1, 12, now))
    s.put(_req(2, 32, now))
    assert [s.get().rid for _ in range(3)] == [1, 2, 0]
    # a large request that has waited long enough goes before a fresh small one
    big = _req(3, 64, now - 1000.0)
    s.put(big)
    s.put(_req(4, 12, now))
    assert s.get() is big


def test_sjf_prefers_a_shape_with_a_resident_context(monkeypatch):
    monkeypatch.setattr(W, "footprint", lambda dims, parms, prec: 200e9)
    resident = {((32, 16, 16, 16), 3, W.MIXED)}
    s = Scheduler("sjf", lambda key: key in resident, create_ms_per_gb=10.0)
    now = time.perf_counter()
    s.put(_req(0, 12, now))   # cheaper, but its context would be created
    s.put(_req(1, 16, now))   # resident
    assert s.get().rid == 1


def test_join_waits_for_task_done():
    s = Scheduler("fifo", lambda key: True)
    s.put(_req(0, 12, time.perf_counter()))
    r = s.get()
    done = []

    def finish():
        time.sleep(0.05)
        done.append(r.rid)
        s.task_done()

    t = threading.Thread(target=finish)
    t.start()
    s.join()
    t.join()
    assert done == [0]


def test_largest_serving_shape_is_admitted_in_both_precisions():
    """64^3x128 (about 19 GB of FP64 links) with APE levels {0,8,24,48}, R,T <= 32:
    the default FP64 layout needs more than 260 GB, the lean one fits."""
    from wloop_amd.serve import context_layout
    budget = 260e9
    r64 = standard_request(0, 64, W.OPEN, W.FP64)
    assert W.footprint(r64.dims, r64.parms, W.FP64) > budget
    lean, nbytes = context_layout(r64, budget)
    assert lean and nbytes <= budget
    rmx = standard_request(1, 64, W.PERIODIC, W.MIXED)
    lean, nbytes = context_layout(rmx, budget)
    assert not lean and nbytes <= budget
    small = standard_request(2, 16, W.OPEN, W.FP64)
    assert context_layout(small, budget) == (False, W.footprint(small.dims, small.parms, W.FP64))
