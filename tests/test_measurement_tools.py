"""CPU checks of the measurement tooling: bench.py's one-GPU contract and the
rocprof summary's mapping of kernel symbols to the traffic keys bench.py reads."""
import os
import subprocess
import sys

import pytest

from wloop_amd import profile_summary as ps
import wloop_amd as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


@pytest.mark.parametrize("symbol,key", [
    ("void wl::contract_stream<4, double, double>(wl::Lat, int)", "1/fp64"),
    ("void wl::contract_chain1<4, float, float, false>(wl::Lat, int)", "3/mixed"),
    ("void wl::contract_chain1<4, float, float, true>(wl::Lat, int)", "3b/mixed"),
    ("void wl::contract_chain1<3, double, double, false>(wl::Lat, int)", "3/fp64"),
    ("void wl::contract_chain2<4, double, double, true>(wl::Lat, int)", "5/fp64"),
    ("void wl::contract_chain<4, 4, 4, float, float>(wl::Lat, int)", "2/mixed"),
    ("void wl::contract_depth<4, 8, float, double>(wl::Lat, int)", "4/mixed"),
    ("void wl::contract_depth<4, 8, double, double>(wl::Lat, int)", "4/fp64"),
    ("void wl::ape_step<float, float>(wl::Lat, float2 const*)", None),
])
def test_kernel_keys(symbol, key):
    assert ps.kernel_key(symbol) == key


def test_traffic_keys_match_bench_kernel_ids():
    # bench.py looks traffic up as f"{kernel id}/{precision}" with the ids of
    # the package; the summary's form numbers must be the same ids
    forms = dict(ps.FORMS)
    assert forms["contract_stream"] == W.KERNEL_STREAM
    assert forms["contract_chain"] == W.KERNEL_CHAIN
    assert forms["contract_chain1"] == W.KERNEL_CHAIN1
    assert forms["contract_depth"] == W.KERNEL_DEPTH
    assert forms["contract_chain2"] == W.KERNEL_CHAIN2
