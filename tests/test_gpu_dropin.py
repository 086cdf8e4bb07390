"""The reference's module entry points (smear.h / wloop.h / lnk.h), called
through the C ABI exactly as ms10 and the devel checks call them, against the
FP64 oracle.  The gauge field is written into udfld() in openQCD's layout
(link (x,mu) at the odd end point, lnkcom.c:840-859) through the even-odd
ipt[] of geometry()."""
import ctypes

import numpy as np
import pytest

import oracle as O
from conftest import rel_err

pytestmark = pytest.mark.gpu


class WloopParms(ctypes.Structure):
    """wloop_parms_t of include/wloop.h"""
    _fields_ = [("msl", ctypes.c_int), ("nss", ctypes.c_int * 5), ("mwlt", ctypes.c_int),
                ("mwlr", ctypes.c_int), ("wls", ctypes.c_int), ("op_smear", ctypes.c_int),
                ("alpha_action", ctypes.c_double * 3), ("alpha_ape", ctypes.c_double),
                ("alpha_hyp", ctypes.c_double * 2), ("proj", ctypes.c_int),
                ("proj_iter", ctypes.c_int), ("wl_mat", ctypes.POINTER(ctypes.c_double))]


@pytest.fixture(scope="module")
def C(wl):
    L = wl.lib()
    L.udfld.restype = ctypes.POINTER(ctypes.c_double)
    L.lnk.restype = ctypes.POINTER(ctypes.c_double)
    L.lnk.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_int]
    L.hyp_links.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int] * 3
    L.hyp_time_links.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int] * 3
    L.ape_spatial_links.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_double] + [ctypes.c_int] * 3
    L.set_wloop_parms.restype = WloopParms
    L.set_wloop_parms.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_double,
                                  ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    L.wloop_parms.restype = WloopParms
    L.wl_idx.restype = ctypes.c_int
    L.get_pus.restype = ctypes.POINTER(ctypes.POINTER(ctypes.c_double))
    L.lnk_plaq_sum.restype = ctypes.c_double
    L.lnk_plaq_sum.argtypes = [ctypes.POINTER(ctypes.c_double)]
    L.alloc_lnk.restype = ctypes.POINTER(ctypes.c_double)
    return L


def setup_lattice(C, dims, bc):
    C.set_lattice_sizes(*dims)
    C.set_bc_type(bc)
    C.geometry()
    V = int(np.prod(dims))
    ipt = np.ctypeslib.as_array(ctypes.cast(ctypes.c_void_p.in_dll(C, "ipt").value,
                                            ctypes.POINTER(ctypes.c_int)), shape=(V,)).copy()
    ud = C.udfld()
    base = ctypes.addressof(ud.contents)
    offs = np.array([[(ctypes.addressof(C.lnk(ud, int(ipt[l]), mu).contents) - base) // 144
                      for mu in range(4)] for l in range(V)])
    return ud, ipt, offs


def write_field(ud, offs, u):
    V = offs.shape[0]
    host = np.ctypeslib.as_array(ud, shape=(4 * V, 18))
    host[offs.reshape(-1)] = u.reshape(-1, 18)


def read_field(ud, offs):
    V = offs.shape[0]
    host = np.ctypeslib.as_array(ud, shape=(4 * V, 18))
    return host[offs.reshape(-1)].reshape(V, 4, 18).copy()


def set_parms(C, nss, mwlt, mwlr, alpha_ape=0.5, proj_iter=4, op=0, alpha_hyp=(0.0, 0.0)):
    n = (ctypes.c_int * 5)(*nss)
    aa = (ctypes.c_double * 3)(1.0, 1.0, 0.5)
    ah = (ctypes.c_double * 2)(*alpha_hyp)
    return C.set_wloop_parms(len(nss), n, mwlt, mwlr, op, aa, alpha_ape, ah, proj_iter)


def test_geometry_is_even_odd(C):
    dims = (4, 4, 6, 4)
    _, ipt, offs = setup_lattice(C, dims, 3)
    V = int(np.prod(dims))
    x = np.indices(dims).reshape(4, -1)
    even = (x.sum(axis=0) % 2) == 0
    assert np.all(ipt[even] < V // 2) and np.all(ipt[~even] >= V // 2)
    assert sorted(offs.reshape(-1).tolist()) == list(range(4 * V))      # a bijection


@pytest.mark.parametrize("bc,precision", [(3, 0), (0, 0), (3, 1)])
def test_ms10_set_data(C, wl, bc, precision):
    """hyp_links(alpha_action) ; wloop_sum(1) ; wl.wl_mat (ms10.c:633-646)."""
    dims = (8, 6, 6, 6)
    nss, T, R = [0, 2, 3], 4, 3
    ud, ipt, offs = setup_lattice(C, dims, bc)
    u = O.generate(dims, bc, 17, 0.4)
    write_field(ud, offs, u)
    C.set_flags(1)
    p = set_parms(C, nss, T, R)
    C.set_wloop_precision(precision)
    C.hyp_links(1.0, 1.0, 0.5, p.proj, p.proj_iter, 0)
    C.wloop_sum(1)
    C.set_wloop_precision(0)
    w = C.wloop_parms()
    got = np.ctypeslib.as_array(w.wl_mat, shape=(w.wls,)).copy()
    ref, uh = O.ms10(dims, bc, u, (1.0, 1.0, 0.5), O.make_parms(nss, T, R))
    assert rel_err(got, ref) < (1e-12 if precision == 0 else 1e-6)
    # hyp_links replaced the links of udfld (hyp_links.c:432-447); the mixed
    # path sums the staples in FP32
    tol = 1e-13 if precision == 0 else 1e-5
    assert np.max(np.abs(read_field(C.udfld(), offs) - uh)) < tol
    assert C.wl_idx(1, 2, 3, 2, 5) == 1 + 3 * (2 + 3 * (3 + 4 * (2 + 3 * 5)))


def test_smear_get_pus(C):
    dims = (4, 6, 6, 6)
    ud, ipt, offs = setup_lattice(C, dims, 3)
    u = O.generate(dims, 3, 5, 0.4)
    write_field(ud, offs, u)
    C.set_flags(1)
    set_parms(C, [1, 2], 2, 2)
    C.smear()
    pus = C.get_pus()
    ref = O.levels(dims, u, O.make_parms([1, 2], 2, 2))
    V = int(np.prod(dims))
    for sl in range(2):
        got = np.array([[np.ctypeslib.as_array(pus[k + 3 * (int(ipt[l]) + V * sl)], shape=(18,))
                         for k in range(3)] for l in range(V)])
        assert np.max(np.abs(got - ref[sl])) < 1e-13
    C.free_smear()


def test_ape_spatial_links_on_lnk_field(C):
    dims = (4, 4, 6, 6)
    ud, ipt, offs = setup_lattice(C, dims, 3)
    u = O.generate(dims, 3, 6, 0.4)
    f = C.alloc_lnk()
    write_field(f, offs, u)
    C.ape_spatial_links(f, 0.5, 1, 4, 0)
    assert np.max(np.abs(read_field(f, offs) - O.ape(dims, u, 0.5, 4))) < 1e-13
    C.free_lnk(f, 0)


@pytest.mark.parametrize("bc", [3, 0])
def test_lnk_plaq_sum(C, bc):
    """lnkplaq.c:68-128: time plaquettes at x0 = N0-1 are left out for open bc."""
    dims = (6, 4, 4, 4)
    ud, ipt, offs = setup_lattice(C, dims, bc)
    u = O.generate(dims, bc, 9, 0.5)
    write_field(ud, offs, u)
    C.set_flags(1)
    got = C.lnk_plaq_sum(C.udfld())
    c = O.to_complex(u).reshape(dims + (4, 3, 3))
    ref = 0.0
    for mu in range(1, 4):
        for nu in range(mu):
            a = c[..., mu, :, :] @ np.roll(c[..., nu, :, :], -1, axis=mu)
            b = c[..., nu, :, :] @ np.roll(c[..., mu, :, :], -1, axis=nu)
            tr = np.real(np.einsum("...ab,...ab->...", np.conj(a), b))
            if nu == 0 and bc == 0:
                tr = tr[:-1]
            ref += tr.sum()
    assert abs(got - ref) < 1e-10 * abs(ref)


def long_links(levels, dims, R):
    """S_R^i(x,k) = F^i_k(x) S_{R-1}^i(x+k,k) with S_1 = F (wloop.c:207-258),
    in numpy from the oracle's levels ([V][3][18] each)."""
    out = []
    for lev in levels:
        f = O.to_complex(lev.reshape(-1, 1, 18)).reshape(dims + (3, 3, 3))
        s = f.copy()
        for _ in range(R - 1):
            s = np.stack([f[..., k, :, :] @ np.roll(s[..., k, :, :], -1, axis=k + 1)
                          for k in range(3)], axis=-3)
        out.append(s)
    return out


def pus_level(C, pus, ipt, V, sl):
    return np.array([[np.ctypeslib.as_array(pus[k + 3 * (int(ipt[l]) + V * sl)], shape=(18,))
                      for k in range(3)] for l in range(V)])


@pytest.mark.parametrize("between", ["plaq", "hyp"])
def test_get_pus_after_other_gpu_work(C, between):
    """Levels of smear() stay those of smear(): a later plaquette or HYP on
    udfld() reuses the device memory they lived in (smear.c:99-127)."""
    dims = (4, 6, 6, 6)
    ud, ipt, offs = setup_lattice(C, dims, 3)
    u = O.generate(dims, 3, 21, 0.4)
    write_field(ud, offs, u)
    C.set_flags(1)
    set_parms(C, [1, 2], 2, 2)
    C.smear()
    if between == "plaq":
        C.lnk_plaq_sum(C.udfld())
    else:
        C.hyp_links(1.0, 1.0, 0.5, 1, 4, 0)
    pus = C.get_pus()
    ref = O.levels(dims, u, O.make_parms([1, 2], 2, 2))
    V = int(np.prod(dims))
    for sl in range(2):
        assert np.max(np.abs(pus_level(C, pus, ipt, V, sl) - ref[sl])) < 1e-13
    C.free_smear()


@pytest.mark.parametrize("keep", [0, 1])
def test_get_pus_after_wloop_sum(C, keep):
    """After wloop_sum the table holds the levels and, in the upper msl slots,
    the spatial transporters of the last R (wloop.c:207-258, :524-608); free_wloop
    leaves them in place (wloop.c:511-521)."""
    dims = (4, 6, 6, 6)
    nss, T, R = [0, 2], 2, 3
    ud, ipt, offs = setup_lattice(C, dims, 3)
    u = O.generate(dims, 3, 22, 0.35)
    write_field(ud, offs, u)
    C.set_flags(1)
    set_parms(C, nss, T, R)
    C.wloop_sum(keep)
    if keep:
        C.free_wloop()
    pus = C.get_pus()
    V = int(np.prod(dims))
    ref = O.levels(dims, u, O.make_parms(nss, T, R))
    for sl in range(2):
        assert np.max(np.abs(pus_level(C, pus, ipt, V, sl) - ref[sl])) < 1e-13
    sr = long_links(ref, dims, R)
    for sl in range(2):
        got = O.to_complex(pus_level(C, pus, ipt, V, 2 + sl).reshape(-1, 1, 18)).reshape(sr[sl].shape)
        assert np.max(np.abs(got - sr[sl])) < 1e-13
    C.free_smear()


@pytest.mark.parametrize("mu,nu", [(1, 2), (0, 3), (3, 0)])
def test_staple_sum_openqcd_layout(C, mu, nu):
    """staple_sum(u1, u2, mu, nu, vb, cbuf): vb(x,mu) += the staples with the
    nu-links from u1 and the mu-links from u2 (staple_sum.c:53-87), fields in
    openQCD's layout; vb's other directions are left as they are."""
    dims = (4, 4, 6, 4)
    ud, ipt, offs = setup_lattice(C, dims, 3)
    u1 = O.generate(dims, 3, 81, 0.5)
    u2 = O.generate(dims, 3, 82, 0.5)
    v0 = O.generate(dims, 3, 83, 0.5)
    fields = []
    for u in (u1, u2, v0):
        f = C.alloc_lnk()
        write_field(f, offs, u)
        fields.append(f)
    C.staple_sum.argtypes = [ctypes.POINTER(ctypes.c_double)] * 2 + [ctypes.c_int] * 2 + \
        [ctypes.POINTER(ctypes.c_double)] * 2
    C.staple_sum(fields[0], fields[1], mu, nu, fields[2], None)
    got = read_field(fields[2], offs)
    want = O.staple_sum(dims, u1, u2, mu, nu, v0)
    assert np.abs(got - want).max() < 1e-13
    for k in range(4):
        if k != mu:
            assert np.array_equal(got[:, k], v0[:, k])
    for f in fields:
        C.free_lnk(f, 0)
