"""wlg_plaquette_noalloc: the plaquette without any device allocation (its
scratch is a static device array), called through the C ABI, equals
wlg_plaquette bit for bit on two calls in a row, the second on a smaller
lattice (the static array keeps the first call's larger contents)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_plaquette_matches(wl):
    L = wl.lib()
    L.wlg_plaquette_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    cases = [((16, 12, 12, 12), wl.PERIODIC, 7, 0.4), ((6, 4, 4, 4), wl.OPEN, 8, 0.5)]
    for dims, bc, seed, eps in cases:
        print(f"noalloc plaquette {dims} bc {bc}", flush=True)
        with wl.Lattice(dims, bc) as lat:
            lat.generate(seed, eps)
            ref = lat.plaquette()
            got = ctypes.c_double()
            rc = L.wlg_plaquette_noalloc(lat._h, ctypes.byref(got))
            assert rc == 0, L.wlg_error()
            assert np.float64(got.value).tobytes() == np.float64(ref).tobytes(), (got.value, ref)
            assert abs(ref) > 0
