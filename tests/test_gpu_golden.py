"""GPU path against the committed golden fixtures of the oracle: the full 8^4
W matrix and, on 24^3x48 (periodic, eps 0.3, APE levels {0,8,24,48}, R,T <= 12),
W at four time slices and summed over x0.  FP64 within 1e-12, mixed within 1e-6."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import rel_err

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 1e-6)])
def test_c1_golden(wl, precision, tol):
    ref = np.load(os.path.join(HERE, "c1_w.npy"))
    p = wl.make_parms([0, 4, 4], 4, 4)
    with wl.Lattice((8, 8, 8, 8), wl.PERIODIC) as lat:
        lat.upload(O.generate((8, 8, 8, 8), O.PERIODIC, 1, 0.4))
        got = lat.measure(p, precision)
    assert rel_err(got, ref) < tol


@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 1e-6)])
def test_c2_golden(wl, precision, tol):
    g = np.load(os.path.join(HERE, "c2_w_sub.npz"))
    dims = (48, 24, 24, 24)
    p = wl.make_parms([0, 8, 16, 24], 12, 12)
    with wl.Lattice(dims, wl.PERIODIC) as lat:
        lat.generate(2, 0.3)          # device generator, same seed as the fixture
        u = lat.download()
        assert abs(np.abs(u).sum() / g["field_sum"] - 1) < 1e-13
        got = lat.measure(p, precision)
    W = O.wl_reshape(got, O.make_parms([0, 8, 16, 24], 12, 12), dims[0])
    print("c2", precision, rel_err(W[list(g["x0"])], g["w"]), rel_err(W.sum(axis=0), g["wsum"]), flush=True)
    assert rel_err(W[list(g["x0"])], g["w"]) < tol
    assert rel_err(W.sum(axis=0), g["wsum"]) < tol
