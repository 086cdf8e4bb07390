"""Pins of the FP64 oracle (CPU only).

The reference (wloop-1.4) cannot be built here (openQCD-1.4 is not vendored)
and ships no golden vectors, so every oracle function is checked against what
the mathematics fixes and against the reference's own invariant tests:
known-answer vectors, fixed points, closed forms, covariance/invariance, and
independent numpy formulations (tests/numpy_ref.py).
"""
import numpy as np
import pytest

import oracle as O
import numpy_ref as N


# ---------------------------------------------------------------- generator

def test_philox_known_answers():
    # Random123 philox4x32-10 known-answer vectors (kat_vectors)
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert O.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert O.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                    [0xa4093822, 0x299f31d0]) == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


@pytest.mark.parametrize("bc", [O.PERIODIC, O.OPEN])
def test_generator_links_are_su3(bc):
    dims = (6, 4, 4, 4)
    u = O.generate(dims, bc, 7, 0.4)
    c = O.to_complex(u)
    n = int(np.prod(dims[1:]))
    live = c if bc == O.PERIODIC else c[:-n]
    eye = np.eye(3)
    assert np.abs(N.dag(live) @ live - eye).max() < 1e-14
    assert np.abs(np.linalg.det(live) - 1).max() < 1e-14
    if bc == O.OPEN:
        assert np.all(u[-n:, 0, :] == 0.0)           # U_0(x0 = N0-1) = 0
        assert np.all(u[-n:, 1:, :] != 0.0)
    assert np.array_equal(u, O.generate(dims, bc, 7, 0.4))     # seeded
    assert not np.array_equal(u, O.generate(dims, bc, 8, 0.4))


def test_generator_gauge_invariant_content():
    # the Haar gauge transform leaves plaquettes at their exp(i eps H) values
    dims = (4, 4, 4, 4)
    u = O.generate(dims, 3, 3, 0.0)                  # eps = 0: pure gauge
    assert abs(O.plaquette(dims, u) / (6 * 256 * 3) - 1) < 1e-14
    u = O.generate(dims, 3, 3, 0.4)
    p = O.plaquette(dims, u) / (6 * 256 * 3)
    assert 0.0 < p < 0.999


def test_haar_moments():
    # Haar SU(3): E|g_ab|^2 = 1/3, E tr g = 0, E |tr g|^2 = 1
    g = O.to_complex(np.array([O.random_su3(5, s, 4) for s in range(20000)]))
    assert abs(np.mean(np.abs(g) ** 2) - 1 / 3) < 0.01
    tr = np.trace(g, axis1=-2, axis2=-1)
    assert abs(np.mean(tr)) < 0.03
    assert abs(np.mean(np.abs(tr) ** 2) - 1) < 0.05


# --------------------------------------------------------------- projection

def _rand_su3(n, seed):
    return O.to_complex(np.array([O.random_su3(seed, s, 4) for s in range(n)]))


def _proj(c, it):
    return O.to_complex(np.array([O.project(O.from_complex(m), it) for m in c]))


def test_projection_fixed_points():
    v = _rand_su3(50, 1)
    for it in (1, 4, 10):
        assert np.abs(_proj(v, it) - v).max() < 1e-14
        assert np.abs(_proj(2.5 * v, it) - v).max() < 1e-14        # scale is removed
    z = np.zeros(18)
    assert np.array_equal(O.project(z, 4), z)
    tiny = np.full(18, 1e-9)                         # sum |u|^2 < sqrt(18) eps: untouched
    assert np.array_equal(O.project(tiny, 4), tiny)


def test_projection_polar_factor():
    # P_inf(V H) = V for Hermitian positive H close to 1 (Newton-Schulz polar step)
    rng = np.random.default_rng(0)
    v = _rand_su3(30, 2)
    a = rng.normal(size=(30, 3, 3)) + 1j * rng.normal(size=(30, 3, 3))
    h = 0.08 * (a + N.dag(a))
    w, q = np.linalg.eigh(h)
    H = q @ (np.exp(w)[..., None] * N.dag(q))
    got = _proj(v @ H, 30)
    assert np.abs(got - v).max() < 1e-12


def test_projection_removes_det_phase():
    # P_inf(V e^{i phi}) = V: the (1 - i Im det X / 3) factor drives det to 1
    v = _rand_su3(30, 3)
    for phi in (0.05, -0.1):
        got = _proj(v * np.exp(1j * phi), 40)
        assert np.abs(got - v).max() < 1e-11


def test_projection_matches_numpy_formulation():
    rng = np.random.default_rng(1)
    y = _rand_su3(40, 4) + 0.3 * (rng.normal(size=(40, 3, 3)) + 1j * rng.normal(size=(40, 3, 3)))
    for it in (0, 1, 4):
        assert np.abs(_proj(y, it) - N.project(y, it)).max() < 1e-13


def test_projection_unitarity_after_30_iterations():
    # devel/smear/check2.c: unitarity and det = 1 after many iterations
    rng = np.random.default_rng(2)
    y = _rand_su3(40, 5) + 0.2 * (rng.normal(size=(40, 3, 3)) + 1j * rng.normal(size=(40, 3, 3)))
    p = _proj(y, 30)
    assert np.abs(N.dag(p) @ p - np.eye(3)).max() < 1e-14
    assert np.abs(np.linalg.det(p) - 1).max() < 1e-14


# ------------------------------------------------------------------ staples

DIMS = (4, 4, 6, 4)


@pytest.mark.parametrize("mu,nu", [(0, 1), (1, 0), (2, 3), (3, 0)])
def test_staple_sum_matches_rolled_numpy(mu, nu):
    a = O.generate(DIMS, 3, 1, 0.5)
    b = O.generate(DIMS, 3, 2, 0.5)
    v = O.generate(DIMS, 3, 3, 0.5)
    got = N.to_grid(O.staple_sum(DIMS, a, b, mu, nu, v), DIMS)
    ref = N.to_grid(v, DIMS).copy()
    ref[..., mu, :, :] += N.staple(N.to_grid(a, DIMS), N.to_grid(b, DIMS), mu, nu)
    assert np.abs(got - ref).max() < 1e-14


def test_staple_gauge_covariance():
    u = O.generate(DIMS, 3, 4, 0.6)
    ug = O.gauge_transform(DIMS, u, 77)
    # recover g(x) from U_mu^g(x) = g(x) U g(x+mu)^dag is not needed: compare
    # gauge-invariant traces of S(x,mu) U_mu(x)^dag
    for mu, nu in [(0, 2), (3, 1)]:
        s = N.to_grid(O.staple_sum(DIMS, u, u, mu, nu), DIMS)[..., mu, :, :]
        sg = N.to_grid(O.staple_sum(DIMS, ug, ug, mu, nu), DIMS)[..., mu, :, :]
        t = np.trace(s @ N.dag(N.to_grid(u, DIMS)[..., mu, :, :]), axis1=-2, axis2=-1)
        tg = np.trace(sg @ N.dag(N.to_grid(ug, DIMS)[..., mu, :, :]), axis1=-2, axis2=-1)
        assert np.abs(t - tg).max() < 1e-13


# --------------------------------------------------------------------- APE

def test_ape_unit_field():
    dims = (4, 4, 4, 4)
    u = O.ape(dims, O.unit_field(dims), 0.5, 0, time=0)       # unprojected
    c = O.to_complex(u)
    assert np.abs(c[:, 1:] - (1 + 4 * 0.5) * np.eye(3)).max() == 0.0
    assert np.abs(c[:, 0] - np.eye(3)).max() == 0.0


@pytest.mark.parametrize("piter", [0, 4])
def test_ape_matches_numpy(piter):
    u = O.generate(DIMS, 3, 5, 0.4)
    got = N.to_grid(O.ape(DIMS, u, 0.5, piter), DIMS)
    ref = N.ape_step(N.to_grid(u, DIMS), 0.5, piter)
    assert np.abs(got - ref).max() < 1e-13


def test_ape4d_matches_numpy():
    u = O.generate(DIMS, 3, 6, 0.4)
    got = N.to_grid(O.ape(DIMS, u, 0.3, 4, time=1), DIMS)
    ref = N.ape_step(N.to_grid(u, DIMS), 0.3, 4, dirs=(0, 1, 2, 3))
    assert np.abs(got - ref).max() < 1e-13


# --------------------------------------------------------------------- HYP

@pytest.mark.parametrize("a,piter", [((1.0, 1.0, 0.5), 4), ((0.75, 0.6, 0.3), 0), ((0.7, 0.5, 0.4), 10)])
def test_hyp_matches_textbook(a, piter):
    dims = (4, 4, 4, 4)
    u = O.generate(dims, 3, 9, 0.4)
    got = N.to_grid(O.hyp(dims, 3, u, *a, piter), dims)
    ref = N.hyp_textbook(N.to_grid(u, dims), *a, piter)
    assert np.abs(got - ref).max() < 1e-12


def test_hyp_time_only_and_space_only():
    dims = (4, 4, 4, 4)
    u = O.generate(dims, 3, 10, 0.4)
    full = O.hyp(dims, 3, u, 1.0, 1.0, 0.5, 4)
    t = O.hyp(dims, 3, u, 1.0, 1.0, 0.5, 4, time=1, space=0)
    s = O.hyp(dims, 3, u, 1.0, 1.0, 0.5, 4, time=0, space=1)
    assert np.array_equal(t[:, 0], full[:, 0]) and np.array_equal(t[:, 1:], u[:, 1:])
    assert np.array_equal(s[:, 1:], full[:, 1:]) and np.array_equal(s[:, 0], u[:, 0])


def test_hyp_equals_scaled_ape_kat():
    # devel/smear/check3.c:95-121: HYP(a,0,0) = (1-a) APE4d(a/(6(1-a))), no projection
    dims = (4, 4, 4, 4)
    u = O.generate(dims, 3, 456, 0.5)
    alpha = 0.5
    h = O.hyp(dims, 3, u, alpha, 0.0, 0.0, 0)
    ap = O.ape(dims, u, alpha / 6.0 / (1 - alpha), 0, time=1)
    assert np.abs(h - (1 - alpha) * ap).max() < 1e-14


def test_hyp_spatial_equals_scaled_ape_kat():
    # devel/smear/check3.c:136-182: HYP3d(a,0) = (1-a) APE3d(a/(4(1-a))) on spatial links
    dims = (4, 4, 4, 4)
    u = O.generate(dims, 3, 457, 0.5)
    alpha = 0.5
    h = O.hyp_spatial(dims, u, alpha, 0.0, 0)
    ap = O.ape(dims, u, alpha / 4.0 / (1 - alpha), 0, time=0)
    assert np.array_equal(h[:, 0], u[:, 0])
    assert np.abs(h[:, 1:] - (1 - alpha) * ap[:, 1:]).max() < 1e-14


def test_hyp_gauge_covariance():
    dims = (4, 4, 4, 4)
    u = O.generate(dims, 3, 11, 0.4)
    ug = O.gauge_transform(dims, u, 5)
    h = O.hyp(dims, 3, u, 1.0, 1.0, 0.5, 4)
    hg = O.hyp(dims, 3, ug, 1.0, 1.0, 0.5, 4)
    # plaquettes of the smeared field are gauge invariant
    assert abs(O.plaquette(dims, h) - O.plaquette(dims, hg)) < 1e-11


# ------------------------------------------------------------- Wilson loops

def test_unit_field_gives_one():
    # devel/wloop/check1.c
    dims = (8, 4, 4, 4)
    p = O.make_parms([0, 1, 2], 4, 2, proj_iter=10)
    w, _ = O.ms10(dims, 3, O.unit_field(dims), (1.0, 1.0, 0.5), p)
    assert np.all(w == 1.0)


def test_unit_field_open_bc_zero_pattern():
    dims = (8, 4, 4, 4)
    u = O.unit_field(dims)
    u[-64:, 0, :] = 0.0
    p = O.make_parms([0, 1], 4, 2, proj_iter=10)
    w, _ = O.ms10(dims, O.OPEN, u, (1.0, 1.0, 0.5), p)
    W = O.wl_reshape(w, p, dims[0])
    for x0 in range(8):
        for T in range(1, 5):
            blk = W[x0, :, T - 1]
            if x0 >= 8 - T:                      # loop crosses x0 = N0-1
                assert np.all(blk == 0.0)
            else:
                assert np.abs(blk - 1).max() < 1e-14


def test_gauge_invariance():
    # devel/wloop/check2.c
    dims = (8, 4, 4, 6)
    u = O.generate(dims, 3, 12345, 0.4)
    p = O.make_parms([0, 1, 3], 4, 3)
    w1, _ = O.ms10(dims, 3, u, (1.0, 1.0, 0.5), p)
    w2, _ = O.ms10(dims, 3, O.gauge_transform(dims, u, 99), (1.0, 1.0, 0.5), p)
    assert np.max(np.abs(w2 / w1 - 1)) < 1e-12


@pytest.mark.parametrize("s", [(1, 0, -1, 0), (-1, 1, 0, 0), (0, 0, 1, 1), (2, 0, 0, -1)])
def test_translation_invariance(s):
    # devel/wloop/check3.c: W(x0) of the shifted field = W(x0 - s0)
    dims = (8, 4, 6, 4)
    u = O.generate(dims, 3, 123456, 0.4)
    p = O.make_parms([0, 2], 3, 2)
    w1, _ = O.ms10(dims, 3, u, (1.0, 1.0, 0.5), p)
    w2, _ = O.ms10(dims, 3, O.shift(dims, u, s), (1.0, 1.0, 0.5), p)
    W1 = O.wl_reshape(w1, p, dims[0])
    W2 = O.wl_reshape(w2, p, dims[0])
    assert np.max(np.abs(W2 - np.roll(W1, s[0], axis=0))) < 1e-13


@pytest.mark.parametrize("bc", [O.PERIODIC, O.OPEN])
def test_wloop_matches_path_walk(bc):
    dims = (6, 4, 4, 4)
    u = O.generate(dims, bc, 21, 0.4)
    p = O.make_parms([0, 1, 2], 3, 2)
    w, uh = O.ms10(dims, bc, u, (1.0, 1.0, 0.5), p)
    lev = O.levels(dims, uh, p)
    levels = [O.to_complex(l).reshape(dims + (3, 3, 3)) for l in lev]
    ref = N.wilson_loops_walk(N.to_grid(uh, dims), levels, 3, 2)
    W = O.wl_reshape(w, p, dims[0])
    assert np.max(np.abs(W - ref)) < 1e-13
    # the level at the source time slice is the first index: W_ij != W_ji
    assert np.max(np.abs(W - np.swapaxes(W, -1, -2))) > 1e-6


def test_abelian_closed_form():
    # constant flux Phi_c in every (0,k) plane, spatial links 1, HYP off:
    # W_ij(R,T) = (1/3) sum_c cos(Phi_c R T) for every level pair and x0
    dims = (8, 6, 6, 6)
    m = np.array([1, 2, -3])
    phi = 2 * np.pi * m / 6
    x = np.indices(dims)
    ph = (x[1] + x[2] + x[3])[..., None] * phi
    u = np.zeros(dims + (4, 3, 3), dtype=complex)
    u[..., 1:, :, :] = np.eye(3)
    u[..., 0, :, :] = np.einsum("...c,cd->...cd", np.exp(1j * ph), np.eye(3))
    uf = N.from_grid(u)
    p = O.make_parms([0, 2], 4, 3)
    w, _ = O.ms10(dims, 3, uf, (0.0, 0.0, 0.0), p)
    W = O.wl_reshape(w, p, dims[0])
    for R in range(1, 4):
        for T in range(1, 5):
            ref = np.mean(np.cos(phi * R * T))
            assert np.abs(W[:, R - 1, T - 1] - ref).max() < 1e-13


def test_open_bc_zero_set():
    dims = (10, 4, 4, 4)
    u = O.generate(dims, O.OPEN, 31, 0.3)
    p = O.make_parms([0, 1], 5, 2)
    w, _ = O.ms10(dims, O.OPEN, u, (1.0, 1.0, 0.5), p)
    W = O.wl_reshape(w, p, dims[0])
    for x0 in range(10):
        for T in range(1, 6):
            blk = W[x0, :, T - 1]
            if x0 >= 10 - T:
                assert np.all(blk == 0.0)
            else:
                assert np.all(blk != 0.0)


def test_reference_loop_structure_is_bitwise_identical():
    dims = (8, 4, 4, 4)
    u = O.generate(dims, 3, 41, 0.4)
    p = O.make_parms([0, 1, 2], 3, 3)
    w0, _ = O.ms10(dims, 3, u, (1.0, 1.0, 0.5), p, mode=0)
    w1, _ = O.ms10(dims, 3, u, (1.0, 1.0, 0.5), p, mode=1)
    assert np.array_equal(w0, w1)


def test_wl_index_layout():
    # wloop_parms.c:431-434
    p = O.make_parms([0, 1, 2], 4, 3)
    assert O.wl_index(p, 1, 2, 3, 2, 5) == 1 + 3 * (2 + 3 * (3 + 4 * (2 + 3 * 5)))


def test_hyp_spatial_abelian_flux_follows_reference_wiring():
    """3-d HYP with alpha2 != 0 on a constant-flux abelian field, no projection:
    every link is U_k times a closed-form factor that depends on which level-1
    links level 2 reads.  Distinct fluxes per plane separate the reference's
    wiring (links decorated in the staple's own plane, hyp_spatial_links.c:
    187-234) from the textbook one (decorated in the orthogonal planes)."""
    dims = (4, 6, 6, 6)
    u, F = N.abelian_flux_field(dims, 1, 2, 3)
    a1, a2 = 0.75, 0.6
    got = O.hyp_spatial(dims, u, a1, a2, 0)
    want = N.scale_links(u, dims, N.hyp3_abelian_factors(F, a1, a2, "reference"))
    other = N.scale_links(u, dims, N.hyp3_abelian_factors(F, a1, a2, "textbook"))
    assert np.abs(got - want).max() < 1e-14
    assert np.abs(got - other).max() > 1e-2
    assert np.array_equal(got[:, 0], u[:, 0])
