"""Writes the golden fixtures of tests/golden/ from the FP64 oracle.

c1_w.npy      W_ij(R,T;x0) of the 8^4 periodic workload (eps 0.4, seed 1, APE levels
              {0,4,8}, R,T <= 4), all 1,152 entries in wl_idx order.
c2_w_sub.npz  24^3x48 periodic workload (eps 0.3, seed 2, APE levels {0,8,24,48},
              R,T <= 12): W at x0 in X0 (as [x0, R-1, T-1, i, j]) and the sum over x0.
c3s_w_sub.npz the precision-form workload at 32^3x64 (open boundaries in time,
              eps 0.25, seed 3, APE levels {0,16,48,96}, R,T <= 16): every code path
              of the 48^3x96 form (fused FP32 blocks T <= 8, FP64 time transporters
              from T = 9, the S_R formation in the chain launches, open-bc zeros)
              at a size whose oracle fits this host (48^3x96 needs about 55 GB in the
              oracle).  W at x0 in (0, 21, 47, 63) -- 47 = N0-1-16 is the last x0
              with every T nonzero, 63 has only zeros -- and the sum over x0.
c4_<L>_w_sub.npz  two members of the 16-configuration batch workload (periodic,
              APE levels {0,8,24}, R,T <= L/2): 16^3x32 (member 0) and 24^3x48
              (member 6), seeds and eps as C4_MEMBERS; W at three x0 and the sum.
The reference itself cannot run here (openQCD is absent); these vectors freeze
the oracle, which is pinned by tests/test_oracle.py.
Usage: python tests/golden/make_golden.py [c1 c2 c3s c4]   (c3s: about 10 minutes on
8 cores and 15 GB of memory; the others a few minutes together)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle as O  # noqa: E402

X0 = (0, 1, 23, 47)
AA = (1.0, 1.0, 0.5)


def c4_members():
    """The 16 lattices of the batch workload: (L, seed, eps), L^3 x 2L periodic,
    6 of L = 16, 6 of L = 24, 4 of L = 32; eps in [0.2, 0.4]."""
    shapes = [16] * 6 + [24] * 6 + [32] * 4
    return [(L, 100 + i, 0.2 + 0.0125 * ((7 * i) % 17)) for i, L in enumerate(shapes)]


def c1():
    dims = (8, 8, 8, 8)
    u = O.generate(dims, O.PERIODIC, 1, 0.4)
    p = O.make_parms([0, 4, 4], 4, 4, alpha_ape=0.5, proj_iter=4)
    w, _ = O.ms10(dims, O.PERIODIC, u, AA, p)
    np.save(os.path.join(HERE, "c1_w.npy"), w)


def sub(name, dims, bc, seed, eps, nss, R, T, x0):
    u = O.generate(dims, bc, seed, eps)
    p = O.make_parms(nss, T, R, alpha_ape=0.5, proj_iter=4)
    w, _ = O.ms10(dims, bc, u, AA, p)
    W = O.wl_reshape(w, p, dims[0])
    np.savez(os.path.join(HERE, name), x0=np.array(x0), w=W[list(x0)], wsum=W.sum(axis=0),
             field_sum=np.abs(u).sum())


def c2():
    sub("c2_w_sub.npz", (48, 24, 24, 24), O.PERIODIC, 2, 0.3, [0, 8, 16, 24], 12, 12, X0)


def c3s():
    sub("c3s_w_sub.npz", (64, 32, 32, 32), O.OPEN, 3, 0.25, [0, 16, 32, 48], 16, 16, (0, 21, 47, 63))


def c4():
    m = c4_members()
    for i in (0, 6):
        L, seed, eps = m[i]
        sub(f"c4_{L}_w_sub.npz", (2 * L, L, L, L), O.PERIODIC, seed, eps, [0, 8, 16], L // 2, L // 2,
            (0, L // 2 - 1, 2 * L - 1))


if __name__ == "__main__":
    targets = sys.argv[1:] or ["c1", "c2", "c3s", "c4"]
    for t in targets:
        print(f"[golden] {t}", flush=True)
        {"c1": c1, "c2": c2, "c3s": c3s, "c4": c4}[t]()
