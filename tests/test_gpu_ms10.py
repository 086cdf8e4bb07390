"""ms10 driver end to end on the GPU: synthetic and exported configurations,
.dat records against the FP64 oracle, -a continuation and the .end flag."""
import ctypes
import os
import struct

import numpy as np
import pytest

import oracle as O
from conftest import rel_err
from ms10_util import MS10, read_dat, run, write_input

pytestmark = pytest.mark.gpu

DIMS, NSS, T, R = (8, 4, 4, 4), (0, 1, 2), 3, 2
WLS = len(NSS) ** 2 * T * R * DIMS[0]


def oracle_w(u, bc):
    w, _ = O.ms10(DIMS, bc, u, (1.0, 1.0, 0.5), O.make_parms(list(NSS), T, R))
    return w


@pytest.mark.parametrize("bc,precision", [(3, 0), (0, 0), (0, 1)])
def test_synthetic_run_and_append(tmp_path, bc, precision):
    d = str(tmp_path)
    write_input(f"{d}/in", d, first=1, last=2, bc=bc, lattice=DIMS, nss=NSS, mwlt=T, mwlr=R,
                synthetic=(7, 0.4), precision=precision)
    r = run(f"{d}/in", d)
    assert r.returncode == 0, (tmp_path / "log" / "run.ms10.log").read_text()
    write_input(f"{d}/in2", d, first=3, last=4, bc=bc, lattice=DIMS, nss=NSS, mwlt=T, mwlr=R,
                synthetic=(7, 0.4), precision=precision)
    r = run(f"{d}/in2", d, "-a")
    assert r.returncode == 0, (tmp_path / "log" / "run.ms10.log").read_text()
    head, recs = read_dat(f"{d}/dat/run.ms10.dat", WLS)
    assert head == (WLS, 3, T, R, DIMS[0])
    assert [nc for nc, _ in recs] == [1, 2, 3, 4]
    tol = 1e-12 if precision == 0 else 1e-6
    for nc, w in recs:
        assert rel_err(w, oracle_w(O.generate(DIMS, bc, 7 + nc, 0.4), bc)) < tol
    log = (tmp_path / "log" / "run.ms10.log").read_text()
    assert "Continuation run" in log and "Configuration no 4 fully processed" in log
    assert (tmp_path / "dat" / "run.ms10.dat~").read_bytes() == (tmp_path / "dat" / "run.ms10.dat").read_bytes()


def test_end_flag_stops_after_current_configuration(tmp_path):
    d = str(tmp_path)
    write_input(f"{d}/in", d, first=1, last=5, lattice=DIMS, nss=NSS, mwlt=T, mwlr=R)
    (tmp_path / "log" / "run.ms10.end").write_text("")
    r = run(f"{d}/in", d)
    assert r.returncode == 0
    _, recs = read_dat(f"{d}/dat/run.ms10.dat", WLS)
    assert [nc for nc, _ in recs] == [1]
    assert not (tmp_path / "log" / "run.ms10.end").exists()
    assert "End flag set, run stopped" in (tmp_path / "log" / "run.ms10.log").read_text()


def read_exported(path, dims):
    """Lexicographic AoS field from the exported format (odd points, 8 links each)."""
    V = int(np.prod(dims))
    with open(path, "rb") as f:
        n = struct.unpack("<4i", f.read(16))
        assert tuple(n) == tuple(dims)
        f.read(8)
        blocks = np.frombuffer(f.read(), dtype="<f8").reshape(V // 2, 8, 18)
    x = np.indices(dims).reshape(4, -1).T
    odd = np.nonzero(x.sum(axis=1) % 2 == 1)[0]
    u = np.zeros((V, 4, 18))
    strides = np.array([dims[1] * dims[2] * dims[3], dims[2] * dims[3], dims[3], 1])
    for b, l in enumerate(odd):
        for mu in range(4):
            u[l, mu] = blocks[b, 2 * mu]                     # U(x, mu)
            y = x[l].copy()
            y[mu] = (y[mu] - 1) % dims[mu]
            u[int(y @ strides), mu] = blocks[b, 2 * mu + 1]  # U(x - mu, mu)
    return u


def test_exported_configurations(tmp_path, wl):
    L = wl.lib()
    d = str(tmp_path)
    os.makedirs(f"{d}/cnfg")
    L.set_lattice_sizes(*DIMS)
    L.set_bc_type(0)
    L.geometry()
    for nc in (1, 2):
        L.random_cnfg(ctypes.c_ulonglong(100 + nc), ctypes.c_double(0.3))
        L.export_cnfg(f"{d}/cnfg/expn{nc}".encode())
    write_input(f"{d}/in", d, name="exp", first=1, last=2, bc=0, lattice=None, nss=NSS, mwlt=T,
                mwlr=R, synthetic=None, cnfg_dir=f"{d}/cnfg")
    r = run(f"{d}/in", d)
    assert r.returncode == 0, (tmp_path / "log" / "exp.ms10.log").read_text()
    _, recs = read_dat(f"{d}/dat/exp.ms10.dat", WLS)
    for nc, w in recs:
        u = read_exported(f"{d}/cnfg/expn{nc}", DIMS)
        assert np.max(np.abs(u - O.generate(DIMS, 0, 100 + nc, 0.3))) < 1e-13
        assert rel_err(w, oracle_w(u, 0)) < 1e-12


def test_imported_configurations_noexp(tmp_path, wl):
    """-noexp: per-process files loc_dir/<name>n<no>_0 in the imported format
    (main/ms10.c:706-713), one process owning the lattice; the log says so."""
    L = wl.lib()
    d = str(tmp_path)
    os.makedirs(f"{d}/loc")
    L.set_lattice_sizes(*DIMS)
    L.set_bc_type(3)
    L.geometry()
    for nc in (2, 4):
        L.random_cnfg(ctypes.c_ulonglong(200 + nc), ctypes.c_double(0.35))
        L.write_cnfg(f"{d}/loc/impn{nc}_0".encode())
    write_input(f"{d}/in", d, name="imp", first=2, last=4, step=2, bc=3, lattice=None, nss=NSS,
                mwlt=T, mwlr=R, synthetic=None)
    txt = open(f"{d}/in").read().replace("[Directories]\n", f"[Directories]\nloc_dir      {d}/loc\n")
    open(f"{d}/in", "w").write(txt)
    r = run(f"{d}/in", d, "-noexp")
    log = (tmp_path / "log" / "imp.ms10.log").read_text()
    assert r.returncode == 0, log
    assert "Configurations are read in imported file format" in log
    _, recs = read_dat(f"{d}/dat/imp.ms10.dat", WLS)
    assert [nc for nc, _ in recs] == [2, 4]
    for nc, w in recs:
        assert rel_err(w, oracle_w(O.generate(DIMS, 3, 200 + nc, 0.35), 3)) < 1e-12
