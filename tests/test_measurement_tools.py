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
    ("void wl::contract_stream<4, double, double>(wl::Lat, int)", "chain/fp64"),
    ("void wl::contract_chain1<4, float, float, false>(wl::Lat, int)", "chain/mixed"),
    ("void wl::contract_chain1<4, float, float, true>(wl::Lat, int)", "chain_build/mixed"),
    ("void wl::contract_chain1<3, double, double, false>(wl::Lat, int)", "chain/fp64"),
    ("void wl::contract_chain2<4, double, double, true, false>(wl::Lat, int)", "chain/fp64"),
    ("void wl::contract_chain2<4, double, double, true, true>(wl::Lat, int)", "chain_build/fp64"),
    ("void wl::contract_chain<4, 4, 4, float, float>(wl::Lat, int)", "chain/mixed"),
    ("void wl::contract_depth<4, 8, float, float, false>(wl::Lat, int)", "fused/mixed"),
    ("void wl::contract_depth<4, 8, float, double, false>(wl::Lat, int)", "fused64/mixed"),
    ("void wl::contract_depth<4, 8, double, double, false>(wl::Lat, int)", "fused/fp64"),
    ("void wl::contract_depth<4, 4, float, float, true>(wl::Lat, int)", "stored/mixed"),
    ("void wl::ape_step<float, float>(wl::Lat, float2 const*)", "ape_step/mixed"),
    ("void wl::ape_step<double, double>(wl::Lat, double2 const*)", "ape_step/fp64"),
    ("void wl::hyp_level2<float, double>(wl::Lat, double2 const*)", "hyp2/mixed"),
    ("wl::finish_wloop(int, int)", "finish/any"),
    ("wl::d2f(long, HIP_vector_type<double, 2u> const*)", None),
])
def test_kernel_keys(symbol, key):
    assert ps.kernel_key(symbol) == key


def test_sq_table_utilisation():
    # 1e9 wave instructions in 2.0 ms: 4e9 SIMD cycles of 1024 x 2.4e9 x 2e-3
    sq = {"ape_step/mixed": {"SQ_INSTS_VALU": 1.0e9, "SQ_INSTS_VALU_FMA_F32": 4.0e8,
                             "SQ_INSTS_VALU_FMA_F64": 2.0e8}}
    _, d = ps.sq_table(sq, {"ape_step/mixed": 2.0})
    assert abs(d["ape_step/mixed"]["valu_issue_util"] - 4e9 / (1024 * 2.4e9 * 2e-3)) < 1e-12
    assert abs(d["ape_step/mixed"]["fma_share"] - 0.6) < 1e-12
