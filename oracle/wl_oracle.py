"""ctypes binding of oracle/liboracle.so (test infrastructure only).

Link fields are numpy float64 arrays of shape (V, 4, 18) in lexicographic site
order (x3 fastest), i.e. the memory image of openQCD's su3_dble per link.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

APE, HYP = 0, 1
OPEN, PERIODIC = 0, 3


class Parms(ctypes.Structure):
    _fields_ = [("msl", ctypes.c_int), ("nss", ctypes.c_int * 5),
                ("mwlt", ctypes.c_int), ("mwlr", ctypes.c_int),
                ("op_smear", ctypes.c_int), ("alpha_ape", ctypes.c_double),
                ("alpha_hyp", ctypes.c_double * 2), ("proj", ctypes.c_int),
                ("proj_iter", ctypes.c_int)]


def make_parms(nss, mwlt, mwlr, alpha_ape=0.5, proj_iter=4, op_smear=APE,
               alpha_hyp=(0.0, 0.0)):
    p = Parms()
    p.msl = len(nss)
    for i, n in enumerate(nss):
        p.nss[i] = n
    p.mwlt, p.mwlr = mwlt, mwlr
    p.op_smear = op_smear
    p.alpha_ape = alpha_ape
    p.alpha_hyp[0], p.alpha_hyp[1] = alpha_hyp
    p.proj = int(proj_iter > 0)
    p.proj_iter = proj_iter
    return p


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


_variants = {}


def lib(variant=None):
    """the parity oracle, or a CPU-baseline build of the same source
    (variant "refflags" or "v3", see oracle/Makefile)"""
    global _lib
    if variant is not None:
        if variant not in _variants:
            path = os.path.join(_HERE, f"liboracle_{variant}.so")
            if not os.path.exists(path):
                build()
            _variants[variant] = _bind(ctypes.CDLL(path))
        return _variants[variant]
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = _bind(ctypes.CDLL(_LIB))
    return _lib


def _bind(L):
    """argument and result types of the oracle's C functions"""
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    L.wo_generate.argtypes = [ip, ctypes.c_int, ctypes.c_uint64, ctypes.c_double, dp]
    L.wo_project.argtypes = [dp, ctypes.c_int]
    L.wo_staple_sum.argtypes = [ip, dp, dp, ctypes.c_int, ctypes.c_int, dp]
    L.wo_hyp.argtypes = [ip, ctypes.c_int, dp] + [ctypes.c_double] * 3 + [ctypes.c_int] * 4
    L.wo_hyp_spatial.argtypes = [ip, dp, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int, ctypes.c_int]
    L.wo_ape.argtypes = [ip, dp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.wo_wloop.argtypes = [ip, dp, ctypes.POINTER(Parms), ctypes.c_int, ctypes.c_int, dp]
    L.wo_ms10.argtypes = [ip, ctypes.c_int, dp, dp, ctypes.POINTER(Parms),
                             ctypes.c_int, ctypes.c_int, dp]
    L.wo_levels.argtypes = [ip, dp, ctypes.POINTER(Parms), dp]
    L.wo_gauge_transform.argtypes = [ip, dp, ctypes.c_uint64]
    L.wo_shift.argtypes = [ip, dp, ip]
    L.wo_plaquette.argtypes = [ip, dp]
    L.wo_plaquette.restype = ctypes.c_double
    L.wo_random_su3.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, dp]
    L.wo_philox4x32.argtypes = [ctypes.POINTER(ctypes.c_uint32)] * 3
    L.wo_set_threads.argtypes = [ctypes.c_int]
    L.wo_time_stages.argtypes = [ip, ctypes.c_int, dp, dp, ctypes.POINTER(Parms),
                                    ctypes.c_int, dp]
    return L


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _dims(dims):
    return (ctypes.c_int * 4)(*dims)


def volume(dims):
    return int(np.prod(dims))


def generate(dims, bc, seed, eps):
    u = np.empty((volume(dims), 4, 18))
    lib().wo_generate(_dims(dims), bc, seed, eps, _dp(u))
    return u


def unit_field(dims):
    u = np.zeros((volume(dims), 4, 18))
    u[:, :, [0, 8, 16]] = 1.0
    return u


def project(m, iters):
    m = np.ascontiguousarray(m, dtype=np.float64).copy()
    lib().wo_project(_dp(m), iters)
    return m


def staple_sum(dims, a, b, mu, nu, v=None):
    v = np.zeros_like(a) if v is None else v.copy()
    lib().wo_staple_sum(_dims(dims), _dp(a), _dp(b), mu, nu, _dp(v))
    return v


def hyp(dims, bc, u, a1, a2, a3, proj_iter, time=1, space=1):
    u = u.copy()
    lib().wo_hyp(_dims(dims), bc, _dp(u), a1, a2, a3, int(proj_iter > 0), proj_iter, time, space)
    return u


def hyp_spatial(dims, u, a1, a2, proj_iter):
    u = u.copy()
    lib().wo_hyp_spatial(_dims(dims), _dp(u), a1, a2, int(proj_iter > 0), proj_iter)
    return u


def ape(dims, u, alpha, proj_iter, time=0):
    u = u.copy()
    lib().wo_ape(_dims(dims), _dp(u), alpha, int(proj_iter > 0), proj_iter, time)
    return u


def wloop(dims, u, parms, mode=1, nthreads=0):
    n = parms.msl ** 2 * parms.mwlt * parms.mwlr * dims[0]
    wl = np.empty(n)
    lib().wo_wloop(_dims(dims), _dp(np.ascontiguousarray(u)), ctypes.byref(parms), mode, nthreads, _dp(wl))
    return wl


def ms10(dims, bc, u, alpha_action, parms, mode=1, nthreads=0):
    """hyp_links(alpha_action) then wloop_sum on a copy of u; returns (W, hyp'd u)."""
    u = u.copy()
    n = parms.msl ** 2 * parms.mwlt * parms.mwlr * dims[0]
    wl = np.empty(n)
    aa = np.asarray(alpha_action, dtype=np.float64)
    lib().wo_ms10(_dims(dims), bc, _dp(u), _dp(aa), ctypes.byref(parms), mode, nthreads, _dp(wl))
    return wl, u


def levels(dims, u, parms):
    lev = np.empty((parms.msl, volume(dims), 3, 18))
    lib().wo_levels(_dims(dims), _dp(np.ascontiguousarray(u)), ctypes.byref(parms), _dp(lev))
    return lev


def gauge_transform(dims, u, seed):
    u = u.copy()
    lib().wo_gauge_transform(_dims(dims), _dp(u), seed)
    return u


def shift(dims, u, s):
    u = u.copy()
    lib().wo_shift(_dims(dims), _dp(u), (ctypes.c_int * 4)(*s))
    return u


def plaquette(dims, u):
    return lib().wo_plaquette(_dims(dims), _dp(np.ascontiguousarray(u)))


def random_su3(seed, site, stream):
    g = np.empty(18)
    lib().wo_random_su3(seed, site, stream, _dp(g))
    return g


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().wo_philox4x32(c, k, o)
    return list(o)


def wl_index(parms, sl1, sl2, t, r, x0):
    return sl1 + parms.msl * (sl2 + parms.msl * (t + parms.mwlt * (r + parms.mwlr * x0)))


def wl_reshape(wl, parms, n0):
    """W as array [x0, r, t, sl2, sl1] -> transposed to [x0, R-1, T-1, sl1, sl2]."""
    a = wl.reshape(n0, parms.mwlr, parms.mwlt, parms.msl, parms.msl)
    return a.transpose(0, 1, 2, 4, 3)


def to_complex(m):
    """(..., 18) float -> (..., 3, 3) complex."""
    m = np.asarray(m)
    return (m[..., 0::2] + 1j * m[..., 1::2]).reshape(m.shape[:-1] + (3, 3))


def from_complex(c):
    c = np.asarray(c)
    out = np.empty(c.shape[:-2] + (18,))
    f = c.reshape(c.shape[:-2] + (9,))
    out[..., 0::2] = f.real
    out[..., 1::2] = f.imag
    return out


def time_stages(dims, bc, u, alpha_action, parms, nthreads=0, variant=None):
    """Seconds of one HYP, one APE step, one L_T update, one S_R update (all
    levels) and one (R,T) contraction of the reference loop structure."""
    t = np.zeros(5)
    aa = np.asarray(alpha_action, dtype=np.float64)
    lib(variant).wo_time_stages(_dims(dims), bc, _dp(np.ascontiguousarray(u)), _dp(aa),
                                ctypes.byref(parms), nthreads, _dp(t))
    return t


def ms10_timed(dims, bc, u, alpha_action, parms, nthreads, variant=None, mode=0):
    """Wall seconds of one full ms10 measurement (HYP + wloop_sum, the reference
    loop structure with mode 0) with a given build and thread count."""
    import time
    u = u.copy()
    n = parms.msl ** 2 * parms.mwlt * parms.mwlr * dims[0]
    wl = np.empty(n)
    aa = np.asarray(alpha_action, dtype=np.float64)
    t0 = time.perf_counter()
    lib(variant).wo_ms10(_dims(dims), bc, _dp(u), _dp(aa), ctypes.byref(parms), mode, nthreads, _dp(wl))
    return time.perf_counter() - t0, wl
