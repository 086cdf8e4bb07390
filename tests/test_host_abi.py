"""Host routines of the C ABI that need no GPU, called through the library as a
C caller calls them: the single-matrix projection (projection.c:120-156) and
the staple helpers of lnkstaple.c:318-410 on fields in openQCD's layout."""
import ctypes

import numpy as np
import pytest

import numpy_ref as N
import oracle as O

DP = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def L(wl):
    lib = wl.lib()
    lib.approx_project_to_su3_dble.argtypes = [DP, ctypes.c_int]
    lib.approx_project_to_su3_dble.restype = None
    lib.lnk.restype = DP
    lib.lnk.argtypes = [DP, ctypes.c_int, ctypes.c_int]
    lib.alloc_lnk.restype = DP
    for f in ("add_up_staple", "add_dn_staple", "add_up_staple_save", "add_dn_staple_save"):
        getattr(lib, f).argtypes = [DP, DP, ctypes.c_int, ctypes.c_int, ctypes.c_int, DP]
        getattr(lib, f).restype = None
    for f in ("add_up_staple_lnks", "add_dn_staple_lnks", "add_up_staple_lnks_save",
              "add_dn_staple_lnks_save"):
        getattr(lib, f).argtypes = [DP, DP, DP, DP]
        getattr(lib, f).restype = None
    return lib


def project_c(L, m, iters):
    m = np.ascontiguousarray(m, dtype=np.float64).copy()
    L.approx_project_to_su3_dble(m.ctypes.data_as(DP), iters)
    return m


@pytest.mark.parametrize("iters", [0, 1, 4, 30])
def test_projection_matches_oracle(L, iters):
    rng = np.random.default_rng(7)
    for _ in range(20):
        m = rng.normal(size=18)
        assert np.abs(project_c(L, m, iters) - O.project(m, iters)).max() < 1e-14


def test_projection_special_cases(L):
    # an SU(3) matrix is a fixed point; a multiple of it projects to it; a
    # matrix below the sqrt(18) eps cutoff is left as it is (projection.c:126-128)
    u = O.generate((2, 2, 2, 2), 3, 3, 0.7)[0, 1]
    assert np.abs(project_c(L, u, 4) - u).max() < 1e-14
    assert np.abs(project_c(L, 2.5 * u, 30) - u).max() < 1e-13
    tiny = np.full(18, 1e-17)
    assert np.array_equal(project_c(L, tiny, 4), tiny)


def openqcd_field(L, dims, u):
    """an alloc_lnk() field holding u ([V][4][18], lexicographic) in openQCD's
    layout, with ipt[] and the link offsets"""
    L.set_lattice_sizes(*dims)
    L.set_bc_type(3)
    L.geometry()
    V = int(np.prod(dims))
    ipt = np.ctypeslib.as_array(ctypes.cast(ctypes.c_void_p.in_dll(L, "ipt").value,
                                            ctypes.POINTER(ctypes.c_int)), shape=(V,)).copy()
    f = L.alloc_lnk()
    base = ctypes.addressof(f.contents)
    offs = np.array([[(ctypes.addressof(L.lnk(f, int(ipt[l]), mu).contents) - base) // 144
                      for mu in range(4)] for l in range(V)])
    np.ctypeslib.as_array(f, shape=(4 * V, 18))[offs.reshape(-1)] = u.reshape(-1, 18)
    return f, ipt


@pytest.mark.parametrize("save", ["", "_save"])
def test_site_staples(L, save):
    """add_up_staple(u1,u2,x,mu,nu,v): v += u1_nu(x) u2_mu(x+nu) u1_nu(x+mu)^dag;
    add_dn_staple: v += u1_nu(x-nu)^dag u2_mu(x-nu) u1_nu(x-nu+mu), with the
    nu-links from u1 and the mu-links from u2 (lnkstaple.c:341-384)"""
    dims = (4, 4, 6, 4)
    u1 = O.generate(dims, 3, 71, 0.5)
    u2 = O.generate(dims, 3, 72, 0.5)
    f1, ipt = openqcd_field(L, dims, u1)
    f2, _ = openqcd_field(L, dims, u2)
    g1, g2 = N.to_grid(u1, dims), N.to_grid(u2, dims)
    up = getattr(L, "add_up_staple" + save)
    dn = getattr(L, "add_dn_staple" + save)
    rng = np.random.default_rng(3)
    for _ in range(12):
        x = tuple(int(rng.integers(n)) for n in dims)
        mu, nu = (int(v) for v in rng.choice(4, size=2, replace=False))
        l = ((x[0] * dims[1] + x[1]) * dims[2] + x[2]) * dims[3] + x[3]
        a1, a2 = g1[..., nu, :, :], g2[..., mu, :, :]
        want_up = N.mm(a1, N.up(a2, nu), N.dag(N.up(a1, mu)))[x]
        want_dn = N.mm(N.dag(N.dn(a1, nu)), N.dn(a2, nu), N.up(N.dn(a1, nu), mu))[x]
        for fn, want in ((up, want_up), (dn, want_dn)):
            v = np.zeros(18)
            fn(f1, f2, int(ipt[l]), mu, nu, v.ctypes.data_as(DP))
            assert np.abs(O.to_complex(v) - want).max() < 1e-14, (fn.__name__, x, mu, nu)


@pytest.mark.parametrize("save", ["", "_save"])
def test_link_staples(L, save):
    """add_up_staple_lnks(u1,u2,u3,v): v += u1 u2 u3^dag;
    add_dn_staple_lnks: v += u1^dag u2 u3 (lnkstaple.c:387-410); v accumulates"""
    m = O.generate((2, 2, 2, 2), 3, 5, 0.8).reshape(-1, 18)[:3].copy()
    c = O.to_complex(m)
    v0 = np.linspace(-1, 1, 18)
    for name, want in (("add_up_staple_lnks", c[0] @ c[1] @ c[2].conj().T),
                       ("add_dn_staple_lnks", c[0].conj().T @ c[1] @ c[2])):
        v = v0.copy()
        getattr(L, name + save)(*(m[i].ctypes.data_as(DP) for i in range(3)), v.ctypes.data_as(DP))
        assert np.abs(O.to_complex(v) - O.to_complex(v0) - want).max() < 1e-14
