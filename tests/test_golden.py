"""The oracle reproduces the committed golden fixtures (tests/golden/)."""
import os

import numpy as np

import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_c1_fixture_bitwise():
    dims = (8, 8, 8, 8)
    u = O.generate(dims, O.PERIODIC, 1, 0.4)
    p = O.make_parms([0, 4, 4], 4, 4, alpha_ape=0.5, proj_iter=4)
    w, _ = O.ms10(dims, O.PERIODIC, u, (1.0, 1.0, 0.5), p)
    assert np.array_equal(w, np.load(os.path.join(HERE, "c1_w.npy")))


def test_c2_fixture_field():
    g = np.load(os.path.join(HERE, "c2_w_sub.npz"))
    u = O.generate((48, 24, 24, 24), O.PERIODIC, 2, 0.3)
    assert np.abs(u).sum() == g["field_sum"]
    assert g["w"].shape == (4, 12, 12, 4, 4) and g["wsum"].shape == (12, 12, 4, 4)
