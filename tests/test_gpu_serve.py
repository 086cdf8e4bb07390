"""The measurement daemon on the GPU: every served W equals the eager
measurement of the same seeded configuration; a small HBM budget forces
context eviction; a request that can never fit is rejected, the rest served."""
import numpy as np
import pytest

from wloop_amd import serve

pytestmark = pytest.mark.gpu


def test_stream_results_and_admission(wl):
    sizes = (4, 6, 8)
    fp = {L: wl.footprint((2 * L, L, L, L), serve.standard_request(0, L, 3, wl.MIXED).parms, wl.MIXED)
          for L in sizes}
    budget = (fp[8] + fp[6]) / 1e9 * 1.01      # forces evictions between shapes
    out, reqs = serve.run_stream(200.0, 24, wl.MIXED, budget, workers=2, sizes=sizes, seed=3)
    print(out, flush=True)
    assert out["completed"] == 24 and out["rejected"] == 0
    assert out["evicted_contexts"] > 0
    assert out["latency_ms"]["p99"] >= out["latency_ms"]["p50"] > 0
    for r in reqs[:6]:
        with wl.Lattice(r.dims, r.bc) as lat:
            lat.generate(r.seed, r.eps)
            assert np.array_equal(lat.measure(r.parms, r.precision), r.result)


def test_oversized_request_rejected(wl):
    srv = serve.Server(int(1e6), workers=1)            # 1 MB budget
    r = serve.standard_request(0, 8, 3, wl.FP64)
    srv.submit(r)
    srv.drain()
    srv.stop()
    assert r.error and "MemoryError" in r.error
