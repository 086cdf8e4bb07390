"""The C host layer, the ms10 driver and the oracle built with
AddressSanitizer and UndefinedBehaviorSanitizer (gcc; the device objects are
the ordinary ones) and run on the paths that need no GPU: every input check of
ms10 and its parameter and data-file records, and every oracle function on
small lattices.  A sanitizer report fails the test."""
import os
import subprocess

import pytest

from ms10_util import write_input

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "wloop_amd", "csrc")
ASAN_MS10 = os.path.join(CSRC, "_build", "asan", "ms10")
ORACLE_CHECK = os.path.join(ROOT, "oracle", "_build", "sanitize_check")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:exitcode=99",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1:exitcode=98", OMP_NUM_THREADS="2")
REPORTS = ("AddressSanitizer", "runtime error:", "LeakSanitizer")


def _make(d):
    r = subprocess.run(["make", "-s", "-C", d, "sanitize"], capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        pytest.skip("sanitizer build failed: " + r.stderr[-500:])


@pytest.fixture(scope="module")
def ms10():
    if not os.path.exists(os.path.join(CSRC, "_build", "api.o")):
        pytest.skip("device objects not built")
    _make(CSRC)
    return ASAN_MS10


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=300, env=ENV)
    for rep in REPORTS:
        assert rep not in r.stderr, r.stderr[-4000:]
    assert r.returncode not in (98, 99), r.stderr[-4000:]
    return r


def test_oracle_functions():
    _make(os.path.join(ROOT, "oracle"))
    r = _run([ORACLE_CHECK], ROOT)
    assert r.returncode == 0 and "sanitize_check: ok" in r.stdout


def test_ms10_syntax(ms10, tmp_path):
    r = _run([ms10], tmp_path)
    assert r.returncode != 0
    assert "Syntax: ms10" in (tmp_path / "STARTUP_ERROR").read_text()


@pytest.mark.parametrize("bc", [0, 1, 2, 3])
def test_ms10_records(ms10, tmp_path, bc):
    d = str(tmp_path)
    write_input(f"{d}/in", d, nss=(0, 2, 3), mwlt=3, mwlr=2, lattice=(8, 4, 4, 4), bc=bc)
    r = _run([ms10, "-i", f"{d}/in"], d)
    if bc in (1, 2):
        assert r.returncode != 0
        assert "Only open (0) and periodic (3)" in (tmp_path / "STARTUP_ERROR").read_text()
        return
    assert (tmp_path / "dat" / "run.ms10.par").stat().st_size > 0
    assert (tmp_path / "dat" / "run.ms10.dat").stat().st_size >= 20


def test_ms10_refusals(ms10, tmp_path):
    d = str(tmp_path)
    write_input(f"{d}/in", d, nss=(0, 1, 2))
    _run([ms10, "-i", f"{d}/in"], d)
    write_input(f"{d}/in2", d, nss=(0, 1, 3), first=3, last=4)
    r = _run([ms10, "-i", f"{d}/in2", "-a"], d)
    assert "Parameters do not match" in (tmp_path / "STARTUP_ERROR").read_text()
    os.remove(tmp_path / "STARTUP_ERROR")
    r = _run([ms10, "-i", f"{d}/in"], d)
    assert r.returncode != 0
    assert "Attempt to overwrite" in (tmp_path / "STARTUP_ERROR").read_text()


@pytest.mark.parametrize("edit,msg", [
    (("mwlt         3", "mwlt         6"), "max wloop time extension out of range"),
    (("op_smear     APE", "op_smear     XYZ"), "Unknown operator smearing XYZ"),
    (("nss          0 1 2", "nss          0 2 1"), ""),
    (("size 8 4 4 4", "size 8 4 4"), ""),
])
def test_ms10_bad_input(ms10, tmp_path, edit, msg):
    d = str(tmp_path)
    write_input(f"{d}/in", d, nss=(0, 1, 2))
    (tmp_path / "in").write_text((tmp_path / "in").read_text().replace(*edit))
    r = _run([ms10, "-i", f"{d}/in"], d)
    assert r.returncode != 0
    if msg:
        assert msg in (tmp_path / "STARTUP_ERROR").read_text()


def test_ms10_noexp_missing_file(ms10, tmp_path):
    d = str(tmp_path)
    os.makedirs(tmp_path / "cnfg")
    write_input(f"{d}/in", d, lattice=None, synthetic=None, cnfg_dir=f"{d}/cnfg")
    r = _run([ms10, "-i", f"{d}/in", "-noexp"], d)
    assert r.returncode != 0
