"""GPU parity of the HIP path (through the C ABI) against the FP64 oracle.

Plain FP64 path: every W_ij(R,T;x0) within 1e-12 relative of the oracle; open
boundary entries that are exactly zero in the oracle are exactly zero here.
Mixed path (FP32 bulk, FP64 projection and sums): within 1e-6 relative.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL_MIXED = 1e-6


def say(*a):
    print(*a, flush=True)


def max_abs(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))))


@pytest.mark.parametrize("dims,bc", [((4, 4, 4, 4), 3), ((6, 4, 4, 4), 0)])
def test_generate_matches_oracle(wl, dims, bc):
    say("generate", dims, bc)
    with wl.Lattice(dims, bc) as lat:
        lat.generate(11, 0.4)
        u = lat.download()
    ref = O.generate(dims, bc, 11, 0.4)
    assert max_abs(u, ref) < 1e-13
    if bc == 0:
        n = int(np.prod(dims[1:]))
        assert np.all(u[-n:, 0, :] == 0.0)


def test_upload_download_roundtrip(wl):
    dims = (4, 4, 6, 8)
    u = O.generate(dims, 3, 5, 0.3)
    with wl.Lattice(dims) as lat:
        lat.upload(u)
        assert np.array_equal(lat.download(), u)


@pytest.mark.parametrize("mu,nu", [(0, 1), (2, 0), (3, 1)])
def test_staple_sum(wl, mu, nu):
    dims = (4, 6, 4, 8)
    a = O.generate(dims, 3, 21, 0.5)
    b = O.generate(dims, 3, 22, 0.5)
    v = O.generate(dims, 3, 23, 0.5)
    ref = O.staple_sum(dims, a, b, mu, nu, v)
    with wl.Lattice(dims) as lat:
        got = lat.staple_sum(a, b, mu, nu, v)
    assert max_abs(got, ref) < 1e-13


@pytest.mark.parametrize("proj_iter", [0, 4, 10])
def test_ape_step(wl, proj_iter):
    dims = (4, 6, 6, 8)
    u = O.generate(dims, 3, 3, 0.4)
    ref = O.ape(dims, u, 0.5, proj_iter, time=0)
    with wl.Lattice(dims) as lat:
        lat.upload(u)
        lat.ape_step(0.5, proj_iter)
        got = lat.download()
    say("ape", proj_iter, max_abs(got, ref))
    assert max_abs(got, ref) < 1e-13


def test_ape_step_4d(wl):
    dims = (4, 4, 6, 6)
    u = O.generate(dims, 3, 4, 0.4)
    ref = O.ape(dims, u, 0.3, 4, time=1)
    with wl.Lattice(dims) as lat:
        lat.upload(u)
        lat.ape_step(0.3, 4, four_d=True)
        got = lat.download()
    assert max_abs(got, ref) < 1e-13


@pytest.mark.parametrize("time,space,bc", [(1, 1, 3), (1, 1, 0), (1, 0, 3), (0, 1, 0)])
def test_hyp(wl, time, space, bc):
    dims = (6, 4, 4, 6)
    u = O.generate(dims, bc, 8, 0.4)
    ref = O.hyp(dims, bc, u, 1.0, 1.0, 0.5, 4, time, space)
    with wl.Lattice(dims, bc) as lat:
        lat.upload(u)
        lat.hyp((1.0, 1.0, 0.5), 4, time=time, space=space)
        got = lat.download()
    say("hyp", time, space, bc, max_abs(got, ref))
    assert max_abs(got, ref) < 1e-13


def test_hyp_spatial_step(wl):
    dims = (4, 6, 6, 6)
    u = O.generate(dims, 3, 9, 0.4)
    ref = O.hyp_spatial(dims, u, 0.75, 0.3, 4)
    with wl.Lattice(dims) as lat:
        lat.upload(u)
        lat.hyp_spatial_step(0.75, 0.3, 4)
        got = lat.download()
    assert max_abs(got, ref) < 1e-13


C1 = dict(dims=(8, 8, 8, 8), bc=3, seed=1, eps=0.4, nss=[0, 4, 4], R=4, T=4)


def run_both(wl, dims, bc, seed, eps, nss, R, T, precision, op=0, alpha_hyp=(0.0, 0.0), proj_iter=4):
    u = O.generate(dims, bc, seed, eps)
    po = O.make_parms(nss, T, R, alpha_ape=0.5, proj_iter=proj_iter, op_smear=op, alpha_hyp=alpha_hyp)
    ref, _ = O.ms10(dims, bc, u, (1.0, 1.0, 0.5), po)
    pg = wl.make_parms(nss, T, R, alpha_ape=0.5, proj_iter=proj_iter, op_smear=op, alpha_hyp=alpha_hyp)
    with wl.Lattice(dims, bc) as lat:
        lat.upload(u)
        got = lat.measure(pg, precision)
    return got, ref


def test_measure_c1_fp64(wl):
    say("C1 fp64")
    got, ref = run_both(wl, precision=wl.FP64, **C1)
    e = rel_err(got, ref)
    say("C1 fp64 max rel err", e)
    assert e < TOL64


def test_measure_open_fp64(wl):
    say("open fp64")
    got, ref = run_both(wl, (12, 6, 6, 6), 0, 2, 0.3, [0, 2, 3], 3, 5, wl.FP64)
    assert np.any(ref == 0.0)
    assert rel_err(got, ref) < TOL64


def test_measure_hyp_operator_fp64(wl):
    got, ref = run_both(wl, (8, 6, 6, 6), 3, 4, 0.35, [0, 1, 2], 3, 3, wl.FP64, op=1,
                        alpha_hyp=(0.75, 0.3))
    assert rel_err(got, ref) < TOL64


def test_measure_c1_mixed(wl):
    say("C1 mixed")
    got, ref = run_both(wl, precision=wl.MIXED, **C1)
    e = rel_err(got, ref)
    say("C1 mixed max rel err", e)
    assert e < TOL_MIXED


def test_measure_open_mixed(wl):
    got, ref = run_both(wl, (20, 8, 8, 8), 0, 3, 0.25, [0, 2, 4], 4, 10, wl.MIXED)
    assert np.any(ref == 0.0)
    e = rel_err(got, ref)
    say("open mixed max rel err", e)
    assert e < TOL_MIXED


def test_levels_match_oracle(wl):
    dims = (4, 6, 6, 6)
    u = O.generate(dims, 3, 12, 0.4)
    po = O.make_parms([1, 2, 0], 2, 2)
    uh = O.hyp(dims, 3, u, 1.0, 1.0, 0.5, 4)
    ref = O.levels(dims, uh, po)
    pg = wl.make_parms([1, 2, 0], 2, 2)
    with wl.Lattice(dims) as lat:
        lat.upload(u)
        lat.measure(pg)
        for sl in range(3):
            assert max_abs(lat.level(sl), ref[sl]) < 1e-13


# level counts at the ends of the reference's range (1 <= msl <= 5,
# wloop_parms.c:156-157; msl = 5 runs the streaming contraction, the other
# forms need msl <= 4)
@pytest.mark.parametrize("nss", [[0], [0, 1, 1, 1, 1]])
@pytest.mark.parametrize("precision", [0, 1])
def test_measure_level_count_edges(wl, nss, precision):
    say("msl", len(nss), "precision", precision)
    got, ref = run_both(wl, (10, 6, 6, 6), 0, 5, 0.3, nss, 3, 5, precision)
    assert rel_err(got, ref) < (TOL64 if precision == 0 else TOL_MIXED)


# without projection (proj_iter = 0: proj = 0, wloop_parms.c:171)
@pytest.mark.parametrize("precision", [0, 1])
def test_measure_without_projection(wl, precision):
    got, ref = run_both(wl, (8, 6, 6, 6), 3, 6, 0.3, [0, 1, 2], 3, 4, precision, proj_iter=0)
    assert rel_err(got, ref) < (TOL64 if precision == 0 else TOL_MIXED)


# mixed path around the FP32/FP64 split: T up to exactly 9 (the chain launch
# that forms S_R is the only chain launch of an R), and R = 1 only (no S_R is
# formed) with T beyond 9
@pytest.mark.parametrize("mwlt,mwlr", [(9, 3), (10, 1)])
def test_measure_mixed_split_edges(wl, mwlt, mwlr):
    say("mixed split edges", mwlt, mwlr)
    got, ref = run_both(wl, (20, 6, 6, 6), 0, 7, 0.3, [0, 2, 3], mwlr, mwlt, wl.MIXED)
    assert rel_err(got, ref) < TOL_MIXED


@pytest.mark.parametrize("precision", [0, 1])
def test_hyp_spatial_abelian_flux(wl, precision):
    """The GPU 3-d HYP step (alpha2 != 0, no projection) on a constant-flux
    abelian field equals the closed form of the reference's level wiring
    (hyp_spatial_links.c:187-234), not the textbook one."""
    import numpy_ref as N
    dims = (4, 6, 6, 6)
    u, F = N.abelian_flux_field(dims, 1, 2, 3)
    a1, a2 = 0.75, 0.6
    with wl.Lattice(dims) as lat:
        lat.upload(u)
        lat.hyp_spatial_step(a1, a2, 0, precision=precision)
        got = lat.download()
    want = N.scale_links(u, dims, N.hyp3_abelian_factors(F, a1, a2, "reference"))
    other = N.scale_links(u, dims, N.hyp3_abelian_factors(F, a1, a2, "textbook"))
    tol = 1e-14 if precision == 0 else 1e-6
    say("hyp3 abelian", precision, max_abs(got, want))
    assert max_abs(got, want) < tol
    assert max_abs(got, other) > 1e-2
