"""wloop_amd -- MI355X-native Wilson-loop measurement path (HYP -> APE -> transporters -> W_ij(R,T)).

Python view of the C ABI of ``libwloop_amd.so`` (``include/wloop_gpu.h``).  The
compute runs in hand-written HIP kernels for gfx950; there is no CPU path: if the
library is missing, loading raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwloop_amd.so")
_lib = None

FP64, MIXED = 0, 1
APE, HYP = 0, 1
OPEN, PERIODIC = 0, 3
T_HYP, T_LEVELS, T_TIME_TR, T_SPACE_TR, T_CONTRACT, T_FINISH = range(6)
(TUNE_FP64_CONTRACT_KERNEL, TUNE_FP64_CONTRACT_WG, TUNE_MIXED_CONTRACT_KERNEL, TUNE_MIXED_CONTRACT_WG,
 TUNE_FP64_CONTRACT_DEPTH, TUNE_FP64_CONTRACT_STARTS, TUNE_MIXED_CONTRACT_DEPTH,
 TUNE_MIXED_CONTRACT_STARTS, TUNE_APE_FORM, TUNE_LEAN, TUNE_MIXED_RESTART, TUNE_FUSED_TRACES) = range(12)
(KERNEL_STREAM, KERNEL_CHAIN, KERNEL_CHAIN1, KERNEL_DEPTH, KERNEL_CHAIN2, KERNEL_HYBRID,
 KERNEL_STORED, KERNEL_STORED9) = 1, 2, 3, 4, 5, 6, 7, 8
KERNEL_NAMES = {KERNEL_STREAM: "streaming", KERNEL_CHAIN: "chain (msl lanes per pair)",
                KERNEL_CHAIN1: "chain (one lane per pair)", KERNEL_DEPTH: "fused time slices",
                KERNEL_CHAIN2: "chain (two lanes per pair)",
                KERNEL_HYBRID: "fused time slices for T <= 8, one-lane chain beyond",
                KERNEL_STORED: "fused time slices for T <= 8, one-lane chain at T = 9, fused over "
                               "the stored transporters from T = 10",
                KERNEL_STORED9: "fused time slices for T <= 8 forming the next R's spatial "
                                "transporters, fused over the stored transporters from T = 9"}
STAGES = ("hyp", "levels", "time_transporter", "space_transporter", "contract", "finish")
L_FUSED, L_CHAIN, L_CHAIN_BUILD, L_FINISH, L_STORED = range(5)
LAUNCH_KINDS = {L_FUSED: "fused", L_CHAIN: "chain", L_CHAIN_BUILD: "chain_build", L_FINISH: "finish",
                L_STORED: "stored"}


class WloopError(RuntimeError):
    pass


class Parms(ctypes.Structure):
    """wlg_parms of include/wloop_gpu.h."""
    _fields_ = [("msl", ctypes.c_int), ("nss", ctypes.c_int * 5),
                ("mwlt", ctypes.c_int), ("mwlr", ctypes.c_int),
                ("op_smear", ctypes.c_int), ("alpha_ape", ctypes.c_double),
                ("alpha_hyp", ctypes.c_double * 2), ("proj_iter", ctypes.c_int),
                ("alpha_action", ctypes.c_double * 3), ("hyp_time_only", ctypes.c_int)]


def make_parms(nss, mwlt, mwlr, alpha_ape=0.5, proj_iter=4, alpha_action=(1.0, 1.0, 0.5),
               op_smear=APE, alpha_hyp=(0.0, 0.0), hyp_time_only=False):
    """nss is the cumulative per-level step list of the [Wloop] section."""
    p = Parms()
    p.msl = len(nss)
    for i, n in enumerate(nss):
        p.nss[i] = n
    p.mwlt, p.mwlr = mwlt, mwlr
    p.op_smear = op_smear
    p.alpha_ape = alpha_ape
    p.alpha_hyp[0], p.alpha_hyp[1] = alpha_hyp
    p.proj_iter = proj_iter
    for i in range(3):
        p.alpha_action[i] = alpha_action[i]
    p.hyp_time_only = int(hyp_time_only)
    return p


def lib():
    """Load libwloop_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise WloopError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                         "(make -C wloop_amd/csrc); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    L.wlg_create.argtypes = [ctypes.POINTER(vp), ip, ctypes.c_int, ctypes.c_int]
    L.wlg_destroy.argtypes = [vp]
    L.wlg_destroy.restype = None
    L.wlg_error.restype = ctypes.c_char_p
    L.wlg_upload.argtypes = [vp, dp]
    L.wlg_download.argtypes = [vp, dp]
    L.wlg_generate.argtypes = [vp, ctypes.c_uint64, ctypes.c_double]
    L.wlg_hyp.argtypes = [vp, dp] + [ctypes.c_int] * 5
    L.wlg_ape_step.argtypes = [vp, ctypes.c_double] + [ctypes.c_int] * 4
    L.wlg_hyp_spatial_step.argtypes = [vp, ctypes.c_double, ctypes.c_double] + [ctypes.c_int] * 3
    L.wlg_set_bc.argtypes = [vp]
    L.wlg_plaquette.argtypes = [vp, dp]
    L.wlg_levels.argtypes = [vp, ctypes.POINTER(Parms), ctypes.c_int]
    L.wlg_trim.argtypes = [vp]
    L.wlg_staple_sum.argtypes = [vp, dp, dp, ctypes.c_int, ctypes.c_int, dp]
    L.wlg_wloop.argtypes = [vp, ctypes.POINTER(Parms), ctypes.c_int, dp]
    L.wlg_measure.argtypes = [vp, ctypes.POINTER(Parms), ctypes.c_int, dp]
    L.wlg_level_download.argtypes = [vp, ctypes.c_int, dp]
    L.wlg_footprint.argtypes = [ip, ctypes.POINTER(Parms), ctypes.c_int]
    L.wlg_footprint.restype = ctypes.c_size_t
    L.wlg_footprint_lean.argtypes = [ip, ctypes.POINTER(Parms), ctypes.c_int]
    L.wlg_footprint_lean.restype = ctypes.c_size_t
    L.wlg_plaquette_noalloc.argtypes = [vp, dp]
    L.wlg_synchronize.argtypes = [vp]
    L.wlg_timers_reset.argtypes = [vp]
    L.wlg_timers_reset.restype = None
    L.wlg_timer_ms.argtypes = [vp, ctypes.c_int]
    L.wlg_timer_ms.restype = ctypes.c_double
    L.wlg_contract_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_long), dp]
    L.wlg_launch_stats.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_long), dp]
    L.wlg_tune.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.wlg_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp),
                                   ctypes.POINTER(Parms), ctypes.c_int]
    L.wlg_batch_run.argtypes = [vp, ctypes.POINTER(dp)]
    L.wlg_batch_ms.argtypes = [vp]
    L.wlg_batch_ms.restype = ctypes.c_double
    L.wlg_batch_destroy.argtypes = [vp]
    L.wlg_batch_destroy.restype = None
    L.wlg_tune_get.argtypes = [vp, ctypes.c_int]
    _lib = L
    return L


def _dp(a):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise WloopError("arrays handed to the library must be C-contiguous float64")
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(rc):
    if rc != 0:
        raise WloopError(lib().wlg_error().decode())


def wls(parms, n0):
    return parms.msl ** 2 * parms.mwlt * parms.mwlr * n0


def footprint(dims, parms, precision=FP64, lean=False):
    """device bytes of one measurement: default layout, or the lean layout
    (TUNE_LEAN = 1)"""
    f = lib().wlg_footprint_lean if lean else lib().wlg_footprint
    return f((ctypes.c_int * 4)(*dims), ctypes.byref(parms), precision)


class Lattice:
    """One lattice resident on one GPU (wlg_ctx)."""

    def __init__(self, dims, bc=PERIODIC, device=0):
        self.dims = tuple(int(d) for d in dims)
        self.bc = bc
        self.volume = int(np.prod(self.dims))
        self._h = ctypes.c_void_p()
        _check(lib().wlg_create(ctypes.byref(self._h), (ctypes.c_int * 4)(*self.dims), bc, device))

    def close(self):
        if self._h:
            lib().wlg_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- field transfer -------------------------------------------------
    def upload(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.shape != (self.volume, 4, 18):
            raise WloopError(f"gauge field must have shape {(self.volume, 4, 18)}, got {u.shape}")
        _check(lib().wlg_upload(self._h, _dp(u)))

    def download(self):
        u = np.empty((self.volume, 4, 18))
        _check(lib().wlg_download(self._h, _dp(u)))
        return u

    def generate(self, seed, eps):
        _check(lib().wlg_generate(self._h, seed, eps))

    # --- smearing -------------------------------------------------------
    def hyp(self, alpha, proj_iter, time=True, space=True, precision=FP64, proj=None):
        a = (ctypes.c_double * 3)(*alpha)
        proj = int(proj_iter > 0) if proj is None else int(proj)
        _check(lib().wlg_hyp(self._h, a, proj, proj_iter, int(time), int(space), precision))

    def ape_step(self, alpha, proj_iter, four_d=False, precision=FP64, proj=None):
        proj = int(proj_iter > 0) if proj is None else int(proj)
        _check(lib().wlg_ape_step(self._h, alpha, proj, proj_iter, int(four_d), precision))

    def hyp_spatial_step(self, a1, a2, proj_iter, precision=FP64, proj=None):
        proj = int(proj_iter > 0) if proj is None else int(proj)
        _check(lib().wlg_hyp_spatial_step(self._h, a1, a2, proj, proj_iter, precision))

    def set_bc(self):
        _check(lib().wlg_set_bc(self._h))

    def plaquette(self):
        out = ctypes.c_double()
        _check(lib().wlg_plaquette(self._h, ctypes.byref(out)))
        return out.value

    def levels(self, parms, precision=FP64):
        _check(lib().wlg_levels(self._h, ctypes.byref(parms), precision))

    def trim(self):
        _check(lib().wlg_trim(self._h))

    def staple_sum(self, a, b, mu, nu, v):
        v = np.ascontiguousarray(v, dtype=np.float64).copy()
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        _check(lib().wlg_staple_sum(self._h, _dp(a), _dp(b), mu, nu, _dp(v)))
        return v

    # --- measurement ----------------------------------------------------
    def wloop(self, parms, precision=FP64):
        out = np.empty(wls(parms, self.dims[0]))
        _check(lib().wlg_wloop(self._h, ctypes.byref(parms), precision, _dp(out)))
        return out

    def measure(self, parms, precision=FP64):
        out = np.empty(wls(parms, self.dims[0]))
        _check(lib().wlg_measure(self._h, ctypes.byref(parms), precision, _dp(out)))
        return out

    def level(self, sl):
        lev = np.empty((self.volume, 3, 18))
        _check(lib().wlg_level_download(self._h, sl, _dp(lev)))
        return lev

    def tune(self, param, value):
        _check(lib().wlg_tune(self._h, param, value))

    def tune_get(self, param):
        return lib().wlg_tune_get(self._h, param)

    def synchronize(self):
        _check(lib().wlg_synchronize(self._h))

    def timers_reset(self):
        lib().wlg_timers_reset(self._h)

    def timers(self):
        return {name: lib().wlg_timer_ms(self._h, i) for i, name in enumerate(STAGES)}

    def contract_stats(self):
        n = ctypes.c_long()
        ms = ctypes.c_double()
        _check(lib().wlg_contract_stats(self._h, ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value

    def launch_stats(self):
        """{kind: (launches, summed ms)} of the contraction launch kinds"""
        out = {}
        for kind, name in LAUNCH_KINDS.items():
            n = ctypes.c_long()
            ms = ctypes.c_double()
            _check(lib().wlg_launch_stats(self._h, kind, ctypes.byref(n), ctypes.byref(ms)))
            out[name] = (n.value, ms.value)
        return out


class Batch:
    """The measurement of several resident lattices (any sizes, one device)
    captured once into one HIP graph (wlg_batch); run() replays it on the
    fields currently resident in the lattices."""

    def __init__(self, lattices, parms, precision=MIXED):
        self.lattices = list(lattices)
        self.parms = list(parms)
        n = len(self.lattices)
        handles = (ctypes.c_void_p * n)(*[l._h for l in self.lattices])
        arr = (Parms * n)(*self.parms)
        self._arr = arr
        self._h = ctypes.c_void_p()
        _check(lib().wlg_batch_create(ctypes.byref(self._h), n, handles, arr, precision))
        self.out = [np.empty(wls(p, l.dims[0])) for l, p in zip(self.lattices, self.parms)]

    def run(self):
        ptrs = (ctypes.POINTER(ctypes.c_double) * len(self.out))(*[_dp(o) for o in self.out])
        _check(lib().wlg_batch_run(self._h, ptrs))
        return [o.copy() for o in self.out]

    def device_ms(self):
        return lib().wlg_batch_ms(self._h)

    def close(self):
        if self._h:
            lib().wlg_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
