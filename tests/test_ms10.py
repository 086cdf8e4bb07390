"""ms10 driver: input handling, safety checks and binary records that are
settled before the first GPU call (these run without a GPU).  Records follow
main/ms10.c:86-156 and modules/wloop/wloop_parms.c:343-382."""
import os
import struct

import pytest

from ms10_util import MS10, run, write_input

pytestmark = pytest.mark.skipif(not os.path.exists(MS10), reason="ms10 not built")


def test_syntax_error(tmp_path):
    r = __import__("subprocess").run([MS10], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode != 0
    assert "Syntax: ms10 -i <input file>" in (tmp_path / "STARTUP_ERROR").read_text()


def test_par_and_dat_header(tmp_path):
    d = str(tmp_path)
    write_input(f"{d}/in", d, nss=(0, 2, 3), mwlt=3, mwlr=2, lattice=(8, 4, 4, 4), bc=0)
    run(f"{d}/in", d)  # without a GPU it stops at the first configuration
    par = (tmp_path / "dat" / "run.ms10.par").read_bytes()
    # bc record: int32 type + 10 doubles; wloop record: msl+4 int32 + 6 doubles
    assert len(par) == 4 + 80 + 4 * (3 + 4) + 48
    assert struct.unpack("<i", par[:4])[0] == 0
    ints = struct.unpack("<7i", par[84:112])
    assert ints == (0, 2, 3, 3, 2, 0, 4)            # nss, mwlt, mwlr, op_smear=APE, proj_iter
    dbl = struct.unpack("<6d", par[112:160])
    assert dbl == (0.5, 0.0, 0.0, 1.0, 1.0, 0.5)    # alpha_ape, alpha_hyp[2], alpha_action[3]
    head = struct.unpack("<5i", (tmp_path / "dat" / "run.ms10.dat").read_bytes()[:20])
    assert head == (9 * 3 * 2 * 8, 3, 3, 2, 8)      # wls, msl, mwlt, mwlr, tmax
    assert (tmp_path / "dat" / "run.ms10.par~").read_bytes() == par
    log = (tmp_path / "log" / "run.ms10.log").read_text()
    assert "8x4x4x4 lattice" in log and "1x1x1x1 process grid" in log
    assert "Open boundary conditions" in log and "nss = 0 2 3" in log


def test_refuses_to_overwrite_log(tmp_path):
    d = str(tmp_path)
    write_input(f"{d}/in", d)
    (tmp_path / "log" / "run.ms10.log").write_text("old")
    r = run(f"{d}/in", d)
    assert r.returncode != 0
    assert "Attempt to overwrite old *.log or *.dat file" in (tmp_path / "STARTUP_ERROR").read_text()


def test_append_needs_matching_parameters(tmp_path):
    d = str(tmp_path)
    write_input(f"{d}/in", d, nss=(0, 1, 2))
    run(f"{d}/in", d)
    write_input(f"{d}/in2", d, nss=(0, 1, 3), first=3, last=4)
    r = run(f"{d}/in2", d, "-a")
    assert r.returncode != 0
    assert "Parameters do not match" in (tmp_path / "STARTUP_ERROR").read_text()


def test_parameter_range_checks(tmp_path):
    d = str(tmp_path)
    write_input(f"{d}/in", d, mwlt=6, lattice=(8, 4, 4, 4))     # mwlt > N0/2+1
    r = run(f"{d}/in", d)
    assert r.returncode != 0
    assert "max wloop time extension out of range" in (tmp_path / "STARTUP_ERROR").read_text()


def test_unknown_smearing(tmp_path):
    d = str(tmp_path)
    write_input(f"{d}/in", d)
    txt = (tmp_path / "in").read_text().replace("op_smear     APE", "op_smear     XYZ")
    (tmp_path / "in").write_text(txt)
    r = run(f"{d}/in", d)
    assert r.returncode != 0
    assert "Unknown operator smearing XYZ" in (tmp_path / "STARTUP_ERROR").read_text()
