"""CPU checks of the C ABI: the library loads and exports every function the
headers in include/ declare (no GPU calls)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = ["smear.h", "wloop.h", "lnk.h", "lattice.h", "utils.h", "wloop_gpu.h"]


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
    names = set()
    for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}]*)\)\s*;", text):
        name = m.group(1)
        if name not in {"if", "for", "while", "return", "sizeof"}:
            names.add(name)
    return names


def exported():
    lib = os.path.join(ROOT, "wloop_amd", "libwloop_amd.so")
    if not os.path.exists(lib):
        pytest.skip("libwloop_amd.so not built")
    out = subprocess.check_output(["nm", "-D", "--defined-only", lib], text=True)
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


@pytest.mark.parametrize("header", HEADERS)
def test_every_declared_function_is_exported(header):
    names = declared(header)
    assert names, header
    missing = names - exported()
    assert not missing, f"{header}: not exported: {sorted(missing)}"


def test_library_loads_without_gpu():
    import wloop_amd
    L = wloop_amd.lib()
    assert L.wlg_footprint is not None
    p = wloop_amd.make_parms([0, 16, 32, 48], 16, 16)
    fp = wloop_amd.footprint((96, 48, 48, 48), p, wloop_amd.MIXED)
    assert 50e9 < fp < 150e9


def test_create_fails_cleanly_without_device():
    import torch  # noqa: F401  (only to learn whether a device exists)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import wloop_amd
    with pytest.raises(wloop_amd.WloopError):
        wloop_amd.Lattice((4, 4, 4, 4))


def test_bad_lattice_rejected():
    import wloop_amd
    with pytest.raises(wloop_amd.WloopError):
        wloop_amd.Lattice((1, 4, 4, 4))
