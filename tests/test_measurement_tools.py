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


@pytest.mark.parametrize("form", [W.KERNEL_STORED, W.KERNEL_STORED9])
def test_kernel_rooflines_per_form(form):
    """algorithmic work per launch kind at 48^3x96: form 8 runs the stored
    launches over T = 9..16 and forms S_{R+1} in the fused launches; form 7
    runs T = 9 as a chain launch (forming S_R for R >= 2) and the stored
    launches over T = 10..16"""
    sys.path.insert(0, ROOT)
    import bench
    parms = W.make_parms([0, 16, 32, 48], 16, 16)
    V = 96 * 48 ** 3
    timers = {"levels": 235.0, "hyp": 31.0}
    if form == W.KERNEL_STORED9:
        lstats = {"fused": (16, 16 * 15.5), "stored": (16, 16 * 15.0), "chain": (0, 0.0),
                  "chain_build": (0, 0.0)}
    else:
        lstats = {"fused": (16, 16 * 11.6), "stored": (16, 16 * 13.6), "chain": (1, 2.0),
                  "chain_build": (15, 15 * 6.0)}
    out = {e["kind"]: e for e in bench.kernel_rooflines("mixed", parms, V, timers, lstats, 1, {}, 4, form)}
    flop_t = V * 3 * (2 * 4 * 198 + 16 * 36)
    t_stored = 8 if form == W.KERNEL_STORED9 else 7
    assert out["stored"]["algorithmic_flop_per_launch"] == pytest.approx(flop_t * t_stored)
    ahead = V * 3 * 4 * 198 * 15 / 16 if form == W.KERNEL_STORED9 else 0
    assert out["fused"]["algorithmic_flop_per_launch"] == pytest.approx(flop_t * 8 + ahead)
    assert ("chain_build" in out) == (form == W.KERNEL_STORED)
    assert out["ape_step"]["launches_per_config"] == 96
