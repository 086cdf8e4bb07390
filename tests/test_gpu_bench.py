"""bench.py on the small 8^4 workload: one JSON line with the fields of its
contract (metric, value, roofline of the dominant kernel, CPU baseline of the
oracle, the host-input rate, the module steps outside the headline path)."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_json_line():
    print("[bench] workload c1, 1 step", flush=True)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--workload", "c1", "--steps", "1",
                        "--warmup", "1", "--batch", "2", "--no-autotune", "--serve-requests", "3",
                        "--serve-sizes", "4,6"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
                "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline",
                "latency_ms", "fp64", "batched_mixed", "host_input", "aux_kernels"):
        assert key in d, key
    assert d["n_gpus"] == 1 and d["steps"] == 1 and d["value"] > 0
    assert d["vs_baseline"] is None
    roof = d["roofline"]
    for key in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert key in roof, key
    assert 0 < roof["frac"] < 1.5 and roof["bound"] in ("hbm", "valu")
    cpu = d["cpu_baseline"]
    assert cpu["kind"] == "port" and cpu["value"] > 0 and cpu["cores"] >= 1 and cpu["sample"]
    assert d["mixed_max_rel_err_vs_fp64"] < 1e-6
    assert d["batched_mixed"]["max_rel_err_vs_fp64"] < 1e-6
    assert set(d["aux_kernels"]) == {"ape_4d_step", "hyp_3d_step", "plaquette"}
    # one roofline entry per launch kind; the primary one is the longest
    kinds = {e["kind"] for e in d["roofline_by_kernel"]["mixed"]}
    assert {"ape_step", "hyp"} <= kinds and ({"fused", "chain"} & kinds)
    assert roof["ms_per_config"] == max(e["ms_per_config"] for e in d["roofline_by_kernel"]["mixed"])
    for e in d["roofline_by_kernel"]["mixed"] + d["roofline_by_kernel"]["fp64"]:
        assert e["avg_launch_ms"] > 0 and e["frac"] > 0
    assert set(cpu["builds"]) == {"reference_flags_O_mno_avx", "O3_x86_64_v3"}
    for name in ("mixed", "fp64"):
        s = d["serving"][name]
        assert s["completed"] == 3 and s["latency_ms"]["p99"] >= s["latency_ms"]["p50"] > 0
