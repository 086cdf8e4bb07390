"""Graph-captured batches: every replay of the captured graph, with eager work
and synchronisations between replays and new fields resident each time,
equals the eager measurement of the same fields bit for bit."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_err

pytestmark = pytest.mark.gpu

SHAPES = [((8, 4, 4, 4), 3, [0, 1, 2], 3, 2), ((12, 6, 6, 6), 0, [0, 2, 4], 5, 3),
          ((8, 8, 8, 8), 3, [0, 4, 4], 4, 4)]


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_replays_match_eager(wl, precision):
    lats = [wl.Lattice(d, bc) for d, bc, *_ in SHAPES]
    parms = [wl.make_parms(nss, T, R) for _, _, nss, T, R in SHAPES]
    batch = wl.Batch(lats, parms, precision)
    for rep in range(3):
        for i, lat in enumerate(lats):
            lat.generate(1000 * rep + i, 0.3 + 0.02 * i)
        got = batch.run()
        assert batch.device_ms() > 0
        for i, lat in enumerate(lats):
            eager = lat.measure(parms[i], precision)        # eager work between replays
            assert np.array_equal(got[i], eager), (rep, i)
    # one replay against the oracle (new fields uploaded from the host)
    for i, (d, bc, nss, T, R) in enumerate(SHAPES):
        lats[i].upload(O.generate(d, bc, 77 + i, 0.35))
    got = batch.run()
    for i, (d, bc, nss, T, R) in enumerate(SHAPES):
        ref, _ = O.ms10(d, bc, O.generate(d, bc, 77 + i, 0.35), (1.0, 1.0, 0.5), O.make_parms(nss, T, R))
        assert rel_err(got[i], ref) < (1e-12 if precision == 0 else 1e-6)
    batch.close()
    for lat in lats:
        lat.close()


def test_batch_rejects_duplicate_context(wl):
    lat = wl.Lattice((8, 4, 4, 4), 3)
    p = wl.make_parms([0, 1], 2, 2)
    with pytest.raises(wl.WloopError):
        wl.Batch([lat, lat], [p, p])
    lat.close()


def test_batched_context_buffers_are_held(wl):
    """While a batch graph uses a context's buffers, nothing may free or move
    them: trim, a tuning change and an eager measurement that needs more memory
    fail; after the batch is destroyed they work again."""
    lat = wl.Lattice((8, 4, 4, 4), 3)
    p = wl.make_parms([0, 1], 2, 2)
    batch = wl.Batch([lat], [p], wl.FP64)
    lat.generate(5, 0.3)
    first = batch.run()[0]
    with pytest.raises(wl.WloopError, match="batch"):
        lat.trim()
    with pytest.raises(wl.WloopError, match="batch"):
        lat.tune(wl.TUNE_FP64_CONTRACT_WG, 2048)
    big = wl.make_parms([0, 1, 2, 3, 4], 5, 3)
    with pytest.raises(wl.WloopError, match="batch"):
        lat.measure(big, wl.FP64)
    assert np.array_equal(lat.measure(p, wl.FP64), first)   # fits: allowed
    assert np.array_equal(batch.run()[0], first)
    batch.close()
    lat.trim()
    lat.tune(wl.TUNE_FP64_CONTRACT_WG, 2048)
    lat.measure(big, wl.FP64)
    lat.close()
