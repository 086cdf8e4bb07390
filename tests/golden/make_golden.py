"""Writes the golden fixtures of tests/golden/ from the FP64 oracle.

c1_w.npy   W_ij(R,T;x0) of the 8^4 periodic workload (eps 0.4, seed 1, APE levels
           {0,4,8}, R,T <= 4), all 1,152 entries in wl_idx order.
c2_w_sub.npz  24^3x48 periodic workload (eps 0.3, seed 2, APE levels {0,8,24,48},
           R,T <= 12): W at x0 in X0 (as [x0, R-1, T-1, i, j]) and the sum over x0.
The reference itself cannot run here (openQCD is absent); these vectors freeze
the oracle, which is pinned by tests/test_oracle.py.
Usage: python tests/golden/make_golden.py   (about two minutes on 8 cores)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle as O  # noqa: E402

X0 = (0, 1, 23, 47)
AA = (1.0, 1.0, 0.5)


def c1():
    dims = (8, 8, 8, 8)
    u = O.generate(dims, O.PERIODIC, 1, 0.4)
    p = O.make_parms([0, 4, 4], 4, 4, alpha_ape=0.5, proj_iter=4)
    w, _ = O.ms10(dims, O.PERIODIC, u, AA, p)
    np.save(os.path.join(HERE, "c1_w.npy"), w)


def c2():
    dims = (48, 24, 24, 24)
    u = O.generate(dims, O.PERIODIC, 2, 0.3)
    p = O.make_parms([0, 8, 16, 24], 12, 12, alpha_ape=0.5, proj_iter=4)
    w, _ = O.ms10(dims, O.PERIODIC, u, AA, p)
    W = O.wl_reshape(w, p, dims[0])
    np.savez(os.path.join(HERE, "c2_w_sub.npz"), x0=np.array(X0), w=W[list(X0)], wsum=W.sum(axis=0),
             field_sum=np.abs(u).sum())


if __name__ == "__main__":
    c1()
    c2()
