"""Admission by HBM footprint (wloop_amd.serve.Admission), on the CPU."""
import threading
import time

import pytest

from wloop_amd.serve import Admission


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
