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
    ("void wl::contract_depth<4, 8, float, float, false, true, true>(wl::Lat, int)", "fused/mixed"),
    ("void wl::contract_depth<4, 8, float, float, true, true, false>(wl::Lat, int)", "stored/mixed"),
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
    # 2 ms at 2.0 GHz: SQ_BUSY_CYCLES = 32 shader engines x 4e6 cycles.  1e9
    # wave instructions: 1e8 on the matrix cores (8e8 busy cycles), 2e8 FP64
    # FMA, 1e8 FP64 adds and 5e7 conversions at 4.9 cycles, 5.5e8 others at 2.5
    sq = {"ape_step/mixed": {"SQ_INSTS_VALU": 1.0e9, "SQ_INSTS_VALU_FMA_F32": 4.0e8,
                             "SQ_INSTS_VALU_FMA_F64": 2.0e8, "SQ_INSTS_VALU_ADD_F64": 1.0e8,
                             "SQ_INSTS_VALU_CVT": 5.0e7, "SQ_INSTS_MFMA": 1.0e8,
                             "SQ_VALU_MFMA_BUSY_CYCLES": 8.0e8, "SQ_BUSY_CYCLES": 32 * 4.0e6}}
    _, d = ps.sq_table(sq, {"ape_step/mixed": 2.0})
    e = d["ape_step/mixed"]
    simd_cycles = 1024 * 4.0e6
    assert e["clock_ghz"] == pytest.approx(2.0)
    assert e["valu_busy"] == pytest.approx((4.9 * 3.5e8 + 2.5 * 5.5e8) / simd_cycles)
    assert e["mfma_busy"] == pytest.approx(8.0e8 / simd_cycles)
    assert e["simd_busy"] == pytest.approx(e["valu_busy"] + e["mfma_busy"])
    assert e["fma_share"] == pytest.approx(6.0e8 / 9.0e8)


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
    # the mixed contraction is FP32 work, the APE step FP32 staples and FP64
    # combination and projection
    assert out["fused"]["peak"] == pytest.approx(bench.VALU_PEAK_FP32_TFLOPS)
    a32, a64 = bench.ape_flop_per_site(4, True)
    assert out["ape_step"]["fp32_flop_per_launch"] == pytest.approx(V * a32)
    assert out["ape_step"]["peak"] == pytest.approx(bench.valu_peak(a32, a64))


def test_valu_peak():
    """time-weighted vector peak of FP32 and FP64 work"""
    sys.path.insert(0, ROOT)
    import bench
    assert bench.valu_peak(1.0, 0.0) == pytest.approx(157.3)
    assert bench.valu_peak(0.0, 1.0) == pytest.approx(78.6)
    # equal work: 2 / (1/157.3 + 1/78.6)
    assert bench.valu_peak(1.0, 1.0) == pytest.approx(2.0 / (1.0 / 157.3 + 1.0 / 78.6))
    # the FP64 path: every kind priced at the FP64 peak; the APE flop split
    # adds up to the same total on both paths
    assert sum(bench.ape_flop_per_site(4, True)) == sum(bench.ape_flop_per_site(4, False))
    assert sum(bench.hyp_flop_per_site(4, True)) == sum(bench.hyp_flop_per_site(4, False))
    parms = W.make_parms([0, 16, 32, 48], 16, 16)
    V = 96 * 48 ** 3
    out = {e["kind"]: e for e in bench.kernel_rooflines(
        "fp64", parms, V, {"levels": 385.0, "hyp": 46.0}, {"fused": (16, 16 * 38.0), "chain": (0, 0.0),
                                                          "chain_build": (0, 0.0)}, 1, {}, 8, W.KERNEL_DEPTH)}
    for kind in ("ape_step", "hyp", "fused"):
        assert out[kind]["peak"] == pytest.approx(78.6)


def test_autotune_candidates_per_path():
    """the FP64 trials skip forms 6 and 7 (the two-lane chain again on that
    path); every path's default shape is one of its candidates"""
    from wloop_amd import autotune
    f64 = autotune.path_candidates(W.FP64)
    mixed = autotune.path_candidates(W.MIXED)
    assert not any(c["kernel"] in (W.KERNEL_HYBRID, W.KERNEL_STORED) for c in f64)
    assert any(c["kernel"] == W.KERNEL_CHAIN2 for c in f64)
    assert mixed == autotune.CANDIDATES
    for prec, cands in ((W.FP64, f64), (W.MIXED, mixed)):
        assert autotune.DEFAULT_SHAPE[prec] in cands
