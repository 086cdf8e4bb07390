"""Launch-shape variants of the contraction give the same W (to summation
order), and the autotuner picks one of its candidates."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
def test_contraction_variants_agree(wl, precision):
    dims, bc = (12, 6, 6, 8), 0
    u = O.generate(dims, bc, 44, 0.35)
    p = wl.make_parms([0, 1, 3, 2], 5, 3)
    ref, _ = O.ms10(dims, bc, u, (1.0, 1.0, 0.5), O.make_parms([0, 1, 3, 2], 5, 3))
    pk = wl.TUNE_MIXED_CONTRACT_KERNEL if precision else wl.TUNE_FP64_CONTRACT_KERNEL
    outs = []
    with wl.Lattice(dims, bc) as lat:
        lat.upload(u)
        for kernel, wgs in ((wl.KERNEL_STREAM, (1, 3, 64)), (wl.KERNEL_CHAIN, (1, 7, 512)),
                            (wl.KERNEL_CHAIN1, (1, 5, 256))):
            for wg in wgs:
                lat.tune(pk, kernel)
                lat.tune(pk + 1, wg)
                outs.append(lat.measure(p, precision))
    tol = 1e-12 if precision == 0 else 1e-6
    for w in outs:
        assert rel_err(w, ref) < tol
        assert rel_err(w, outs[0]) < (1e-13 if precision == 0 else 1e-7)


def test_autotune_picks_a_candidate(wl, tmp_path):
    from wloop_amd import autotune
    dims = (8, 8, 8, 8)
    p = wl.make_parms([0, 4, 4], 4, 4)
    with wl.Lattice(dims, 3) as lat:
        lat.generate(1, 0.4)
        cache = str(tmp_path / "tune.json")
        ch = autotune.tune(lat, p, wl.FP64, cache=cache, candidates={wl.KERNEL_STREAM: (2, 4),
                                                                     wl.KERNEL_CHAIN: (64,)})
        assert (ch["kernel"], ch["wg"]) in {(1, 2), (1, 4), (2, 64)}
        assert lat.tune_get(wl.TUNE_FP64_CONTRACT_KERNEL) == ch["kernel"]
        again = autotune.tune(lat, p, wl.FP64, cache=cache)
        assert again["cached"] and again["kernel"] == ch["kernel"]
