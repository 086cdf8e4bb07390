#!/usr/bin/env python3
"""Benchmark of the Wilson-loop measurement path on one MI355X.

Workload (one "step" = one configuration measured end to end):
    48^3 x 96 lattice, open boundaries in time, seeded synthetic links
    U = exp(i eps H) (eps = 0.25) under a Haar gauge transformation;
    HYP of the static action (alpha = 1.0, 1.0, 0.5), APE operator levels
    {0,16,48,96} (nss = 0 16 32 48, alpha = 0.5, proj_iter = 4),
    W_ij(R,T;x0) for R,T = 1..16.
The field is generated on the device before timing, so inputs are resident in
HBM when the clock starts; each step runs HYP -> levels -> transporters ->
contraction and copies W (3.1 MB) to the host.

Paths: "mixed" (FP32 staples and transporter products, FP64 projection, FP64
time-transporter chain beyond T = 8, FP64 sums) is the headline value; the
plain FP64 path is timed beside it, and the mixed W is compared with it.

Reported beside the value: one roofline entry per launch kind of each path
("roofline_by_kernel"; "roofline" is the kind that takes the most time of the
mixed path), the CPU oracle on this host per build and thread count
("cpu_baseline"), the upload through the drop-in boundary ("host_input"), a
short serving stream in both precisions ("serving") and the module steps
outside the headline path ("aux_kernels").  --workload c4 measures the
16-lattice batch instead; --cpu-baseline-full times whole 8^4 and 24^3x48
oracle measurements.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# MI355X vector peaks.  FP32 arithmetic in its 4-byte (VOP2) forms issues at
# twice the FP64 rate on gfx950 without packed math (v_fmac/v_mul/v_add_f32
# ~2 cycles per wave64 instruction, v_fma_f64 ~4.5), so FP32 work is priced
# against the FP32 vector peak; a launch doing both is priced against the
# time-weighted peak F / (F32/P32 + F64/P64)
VALU_PEAK_FP64_TFLOPS = 78.6
VALU_PEAK_FP32_TFLOPS = 157.3


def valu_peak(f32, f64):
    """the vector peak of f32 FP32 and f64 FP64 flop done together"""
    return (f32 + f64) / (f32 / VALU_PEAK_FP32_TFLOPS + f64 / VALU_PEAK_FP64_TFLOPS)

WORKLOADS = {
    "c3": dict(dims=(96, 48, 48, 48), bc=0, eps=0.25, seed=3, nss=[0, 16, 32, 48], R=16, T=16,
               name="48^3x96 open bc, eps=0.25, HYP(1,1,0.5) + APE levels {0,16,48,96}, R,T<=16"),
    "c2": dict(dims=(48, 24, 24, 24), bc=3, eps=0.3, seed=2, nss=[0, 8, 16, 24], R=12, T=12,
               name="24^3x48 periodic, eps=0.3, HYP(1,1,0.5) + APE levels {0,8,24,48}, R,T<=12"),
    "c1": dict(dims=(8, 8, 8, 8), bc=3, eps=0.4, seed=1, nss=[0, 4, 4], R=4, T=4,
               name="8^4 periodic, eps=0.4, HYP(1,1,0.5) + APE levels {0,4,8}, R,T<=4"),
}
ALPHA_ACTION = (1.0, 1.0, 0.5)
ALPHA_APE = 0.5
PROJ_ITER = 4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_err(got, ref):
    zero = ref == 0.0
    if np.any(got[zero] != 0.0):
        return float("inf")
    nz = ~zero
    return float(np.max(np.abs(got[nz] - ref[nz]) / np.abs(ref[nz]))) if nz.any() else 0.0


def time_path(lat, parms, precision, steps, warmup):
    for _ in range(warmup):
        lat.measure(parms, precision)
    lat.synchronize()
    lat.timers_reset()
    lat_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        w = lat.measure(parms, precision)
        lat_ms.append((time.perf_counter() - ts) * 1e3)
    lat.synchronize()
    total = time.perf_counter() - t0
    return w, total, lat_ms, lat.timers(), lat.launch_stats()


def hyp_flop_per_site(p, mixed=False):
    """(FP32, FP64) flop per site: 48 staple sums (4 products and the
    accumulation each; FP32 on the mixed path) and 28 combinations
    (1-a)U + bS with the projection (55 + 525 p flop; always FP64)"""
    staples, comb = 48 * (4 * 198 + 35), 28 * (56 + 55 + 525 * p)
    return (staples, comb) if mixed else (0, staples + comb)


def ape_flop_per_site(p, mixed=False):
    """(FP32, FP64) flop per site, 3 spatial links: 8 staple products and 54
    flop of their accumulation (FP32 on the mixed path), 54 flop of the
    combination and the projection (always FP64)"""
    staples, comb = 3 * (8 * 198 + 54), 3 * (54 + 55 + 525 * p)
    return (staples, comb) if mixed else (0, staples + comb)


KERNEL_LABELS = {
    ("ape_step", "mixed"): "ape_step<float, float, true> (packed FP32 links)",
    ("ape_step", "fp64"): "ape_step<double, double, false>",
    ("hyp", "mixed"): "hyp_level1/2/3<float, double, float> (3 launches)",
    ("hyp", "fp64"): "hyp_level1/2/3<double, double, double> (3 launches)",
    ("fused", "mixed"): "contract_depth<4, {W}, float, float, false, true> (fused time slices, FP32 blocks; "
                        "traces on the matrix cores)",
    ("fused", "fp64"): "contract_depth<4, {W}, double, double, false, true> (fused time slices; traces on the "
                       "matrix cores)",
    ("stored", "mixed"): "contract_depth<4, {W}, float, float, true, true> (fused over the stored transporters; "
                         "traces on the matrix cores)",
    ("chain", "mixed"): "contract_chain1<4, float, float, false> (one-lane chain)",
    ("chain_build", "mixed"): "contract_chain1<4, float, float, true> (one-lane chain forming S_R)",
    ("chain", "fp64"): "contract_chain2<4, double, double, *, false> (two-lane chain)",
    ("chain_build", "fp64"): "contract_chain2<4, double, double, *, true> (two-lane chain forming S_R)",
}


def kernel_rooflines(prec, parms, V, timers, lstats, steps, profile, starts=8, form=None):
    """One roofline entry per launch kind of a path: the APE step, HYP and the
    fused time-slice launches against the vector FLOP peak of their FP32 and
    FP64 work (valu_peak; on the mixed path the staple sums and the
    contraction are FP32, the combinations and projections FP64), the
    per-(R,T) chain launches and the chain launch that forms S_R against
    HBM.  achieved = algorithmic work per launch / average launch time (HIP
    events on the context's stream around each group of launches; the APE step
    from the levels stage over its step count, HYP from its stage per
    configuration); traffic = the HBM bytes per launch of the committed rocprof
    summary, where it has the kind.  starts: start slices per workgroup of
    the fused launches (their template parameter W); form: the contraction
    form the path ran (W.KERNEL_*): form 8 runs the stored blocks from T = 9
    (form 7 from T = 10), and in forms 4 and 8 the first fused launch of each
    R also forms the next R's spatial transporters (one product per level
    link, counted in the fused launches' work)."""
    import wloop_amd as W
    link = 72 if prec == "mixed" else 144
    msl = parms.msl
    nsteps = sum(parms.nss[:msl])
    pairs = parms.mwlr * parms.mwlt
    traffic = profile.get("bytes_per_launch", {})
    out = []

    def entry(kind, per_launch, unit_kind, launches_per_cfg, ms_per_launch, key=None):
        """per_launch: bytes, or (FP32 flop, FP64 flop)"""
        if ms_per_launch <= 0:
            return
        if unit_kind == "flop":
            f32, f64 = per_launch
            per_launch = f32 + f64
            ach = per_launch / (ms_per_launch * 1e-3) / 1e12
            peak = valu_peak(f32, f64)
            e = {"bound": "valu", "achieved": ach, "peak": peak, "unit": "TFLOP/s",
                 "frac": ach / peak, "algorithmic_flop_per_launch": per_launch,
                 "fp32_flop_per_launch": f32, "fp64_flop_per_launch": f64}
        else:
            ach = per_launch / (ms_per_launch * 1e-3) / 1e9
            e = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": ach / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": per_launch}
        tb = traffic.get(key or f"{kind}/{prec}")
        e.update({"traffic": tb, "kernel": KERNEL_LABELS.get((kind, prec), kind).format(W=starts), "kind": kind,
                  "avg_launch_ms": ms_per_launch, "launches_per_config": launches_per_cfg,
                  "ms_per_config": ms_per_launch * launches_per_cfg})
        if tb:
            e["measured_hbm_GBps"] = tb / (ms_per_launch * 1e-3) / 1e9
        out.append(e)

    mixed = prec == "mixed"

    def flop(f, scale=1.0):  # (FP32, FP64) flop: the mixed path's contraction is FP32
        return (f * scale, 0.0) if mixed else (0.0, f * scale)

    if nsteps:
        a32, a64 = ape_flop_per_site(parms.proj_iter, mixed)
        entry("ape_step", (V * a32, V * a64), "flop", nsteps, timers["levels"] / steps / nsteps)
    h32, h64 = hyp_flop_per_site(parms.proj_iter, mixed)
    entry("hyp", (V * h32, V * h64), "flop", 1, timers["hyp"] / steps)
    n_chain = lstats["chain"][0] / steps
    n_build = lstats["chain_build"][0] / steps
    n_fused = lstats["fused"][0] / steps
    n_stored = lstats.get("stored", (0, 0.0))[0] / steps
    # T values per launch: the stored blocks cover T = 10 (form 8: T = 9) ..
    # mwlt of every R
    first_stored = 9 if form == W.KERNEL_STORED9 else 10
    t_stored = max(parms.mwlt - first_stored + 1, 0) * parms.mwlr / n_stored if n_stored else 0
    flop_per_t = V * 3 * (2 * msl * 198 + msl * msl * 36)
    if n_fused:
        t_per_launch = (pairs - n_chain - n_build - n_stored * t_stored) / n_fused
        ahead = V * 3 * msl * 198 * (parms.mwlr - 1) / n_fused if form in (W.KERNEL_DEPTH, W.KERNEL_STORED9) else 0
        entry("fused", flop(flop_per_t * t_per_launch + ahead), "flop", n_fused,
              lstats["fused"][1] / lstats["fused"][0])
    if n_stored:
        entry("stored", flop(flop_per_t * t_stored), "flop", n_stored, lstats["stored"][1] / lstats["stored"][0])
    if n_chain:
        entry("chain", V * (1 + 3 * msl) * link, "bytes", n_chain, lstats["chain"][1] / lstats["chain"][0])
    if n_build:
        # L_T, F and S_{R-1} read, S_R written (the sinks are formed, not read)
        entry("chain_build", V * (1 + 9 * msl) * link, "bytes", n_build,
              lstats["chain_build"][1] / lstats["chain_build"][0])
    return out


def host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))  # the host's CPU share for this job
    return model, n


def cpu_baseline(wl_cfg, parms_o):
    """The oracle (the reference loop structure, FP64) timed on this host, per
    build and thread count, on bounded slabs of the workload (one HYP, one APE
    step, one L_T and one S_R update, one (R,T) contraction each), extrapolated
    per site and per stage count to the whole configuration:
      reference flags (-O -mno-avx, as main/Makefile:86-88), 1 thread: 24^3x4 slab;
      reference flags, all host threads: 48^3x32 slab;
      -O3 -march=x86-64-v3, 1 thread: 24^3x4 slab;
      -O3 -march=x86-64-v3, all host threads: 48^3x32 slab.
    The oracle spreads its threads over time slices, so the all-thread slabs
    have at least as many slices as threads.  About 30-60 s of CPU work."""
    import oracle as O
    dims = wl_cfg["dims"]
    model, ncpu = host_cpu()
    steps = sum(wl_cfg["nss"])
    R, T = wl_cfg["R"], wl_cfg["T"]
    out = {}
    for variant, label in (("refflags", "reference_flags_O_mno_avx"), ("v3", "O3_x86_64_v3")):
        res = {}
        for threads, sdims in ((1, (4,) + (min(24, dims[1]),) * 3), (ncpu, (min(max(32, ncpu), dims[0]),) + tuple(dims[1:]))):
            u = O.generate(sdims, 3, wl_cfg["seed"], wl_cfg["eps"])
            t = O.time_stages(sdims, 3, u, ALPHA_ACTION, parms_o, threads, variant=variant)
            f = float(np.prod(dims)) / float(np.prod(sdims))
            t_cfg = f * (t[0] + steps * t[1] + R * (T - 1) * t[2] + (R - 1) * t[3] + R * T * t[4])
            res[f"threads_{threads}"] = {"configs_per_s": 1.0 / t_cfg, "seconds_per_config": t_cfg,
                                         "slab": "x".join(map(str, sdims[::-1])), "sample_s": float(t.sum())}
            log(f"[bench] cpu baseline {label} {threads} threads: {t_cfg:.1f} s per config ({t.sum():.1f} s sample)")
        out[label] = res
    best = out["O3_x86_64_v3"][f"threads_{ncpu}"]
    return {"value": best["configs_per_s"], "unit": "configs/s", "cores": ncpu, "kind": "port",
            "sample": (f"oracle stages (1 HYP, 1 APE step, 1 L_T update, 1 S_R update, 1 (R,T) "
                       f"contraction) on a {best['slab']} slab, -O3 -march=x86-64-v3 build, {ncpu} "
                       f"threads, extrapolated per site and stage count ({steps} APE steps, "
                       f"{R}x{T - 1} L_T, {R - 1} S_R, {R * T} contractions); the other builds and "
                       f"thread counts in 'builds'"),
            "seconds_per_config": best["seconds_per_config"], "host_cpu": model, "builds": out}


def cpu_baseline_full(log_out):
    """Whole ms10 measurements of the 8^4 and 24^3x48 workloads by the oracle
    (reference loop structure), per build and thread count; the oracle reports
    the end of its stages on stderr meanwhile (a one-thread 24^3x48 run takes
    minutes)."""
    import oracle as O
    os.environ["WO_PROGRESS"] = "1"
    model, ncpu = host_cpu()
    res = {"host_cpu": model, "threads_all": ncpu, "runs": []}
    for name in ("c1", "c2"):
        cfg = WORKLOADS[name]
        p = O.make_parms(cfg["nss"], cfg["T"], cfg["R"], alpha_ape=ALPHA_APE, proj_iter=PROJ_ITER)
        u = O.generate(cfg["dims"], cfg["bc"], cfg["seed"], cfg["eps"])
        for variant in ("refflags", "v3"):
            for threads in (1, ncpu):
                log_out(f"[bench] cpu full {name} {variant} {threads} threads: start")
                t, _ = O.ms10_timed(cfg["dims"], cfg["bc"], u, ALPHA_ACTION, p, threads, variant=variant)
                r = {"workload": cfg["name"], "build": variant, "threads": threads, "seconds": t,
                     "configs_per_s": 1.0 / t}
                log_out(f"[bench] cpu full {name} {variant} {threads} threads: {t:.2f} s")
                res["runs"].append(r)
    return res


# flop per site of the smearing steps outside the headline path (M = 198 flop per
# 3x3 complex product, projection with proj_iter = 4: 2,155 flop, lincomb and
# staple accumulation 108 flop per link): a 4-d APE step (6 staples per link, 4
# links), a 3-d HYP step (level 1: 6 links of 2 staples, level 2: 3 links of 4)
APE4D_FLOP_PER_SITE = 4 * (6 * 2 * 198 + 108 + 2155)
HYP3D_FLOP_PER_SITE = 6 * (2 * 2 * 198 + 108 + 2155) + 3 * (4 * 2 * 198 + 108 + 2155)


def aux_kernels(lat, steps):
    """Steps of the module entry points the headline path does not use, timed on
    the resident field (which they modify, so this runs last): 4-d APE
    (ape_links), 3-d HYP (hyp_spatial_links, operator smearing HYP), and the
    plaquette (lnk_plaq_sum).  ms per step of the entry point (the smearing
    kernels plus the copy of the result into the resident field, which the
    reference's modules update in place) and the fraction of the FP64 vector
    peak of their flop per site."""
    import wloop_amd as W
    V = float(np.prod(lat.dims))
    out = {}
    for name, fn, flop in (
            ("ape_4d_step", lambda: lat.ape_step(ALPHA_APE, PROJ_ITER, four_d=True), APE4D_FLOP_PER_SITE),
            ("hyp_3d_step", lambda: lat.hyp_spatial_step(0.75, 0.6, PROJ_ITER), HYP3D_FLOP_PER_SITE),
            ("plaquette", lambda: lat.plaquette(), None)):
        fn()
        lat.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        lat.synchronize()
        ms = (time.perf_counter() - t0) / steps * 1e3
        d = {"ms_per_step": ms, "precision": "fp64"}
        if flop:
            tf = flop * V / (ms * 1e-3) / 1e12
            d.update({"flop_per_site": flop, "achieved_tflops": tf, "frac_of_fp64_vector_peak": tf / VALU_PEAK_FP64_TFLOPS})
        out[name] = d
        log(f"[bench] {name}: {d}")
    return out


def serving_leg(n, sizes):
    """The serving form: a Poisson stream (mean 4 requests/s) of n seeded
    configurations per precision, sizes uniform over 12^3x24 .. 64^3x128, half
    open and half periodic, each HYP + APE levels {0,8,24,48} + W up to L/2,
    admitted by HBM footprint under a 260 GB budget (the lean layout where the
    default one does not fit); configs/s and p50/p99 latency per precision."""
    import wloop_amd as W
    from wloop_amd import serve
    out = {}
    for prec, name in ((W.MIXED, "mixed"), (W.FP64, "fp64")):
        res, _ = serve.run_stream(4.0, n, prec, 260.0, workers=2, sizes=sizes, seed=21, log=log)
        out[name] = res
        log(f"[bench] serving {name}: {res}")
    return out


def batch_workload(args):
    """16 independent periodic configurations of mixed sizes (6x 16^3x32,
    6x 24^3x48, 4x 32^3x64; eps uniform in [0.2,0.4]), APE levels {0,8,24},
    R,T up to L/2, mixed path: one captured graph vs the same measurements
    one after the other."""
    import random
    import wloop_amd as W
    rng = random.Random(4)
    shapes = [16] * 6 + [24] * 6 + [32] * 4
    lats, parms, seeds, epss = [], [], [], []
    for i, L in enumerate(shapes):
        lats.append(W.Lattice((2 * L, L, L, L), W.PERIODIC))
        parms.append(W.make_parms([0, 8, 16], L // 2, L // 2, alpha_ape=ALPHA_APE, proj_iter=PROJ_ITER,
                                  alpha_action=ALPHA_ACTION))
        seeds.append(100 + i)
        epss.append(rng.uniform(0.2, 0.4))

    def refresh(rep):
        for i, l in enumerate(lats):
            l.generate(seeds[i] + 1000 * rep, epss[i])

    refresh(0)
    seq = []
    for rep in range(args.warmup + args.steps):
        refresh(rep)
        for l in lats:
            l.synchronize()
        t0 = time.perf_counter()
        outs_seq = [l.measure(p, W.MIXED) for l, p in zip(lats, parms)]
        if rep >= args.warmup:
            seq.append(time.perf_counter() - t0)
    batch = W.Batch(lats, parms, W.MIXED)
    gr = []
    for rep in range(args.warmup + args.steps):
        refresh(rep)
        for l in lats:
            l.synchronize()
        t0 = time.perf_counter()
        outs = batch.run()
        if rep >= args.warmup:
            gr.append(time.perf_counter() - t0)
    same = all(np.array_equal(a, b) for a, b in zip(outs, outs_seq))
    batch.close()
    n = len(lats)
    out = {"metric": "configs/s, 16-configuration mixed-size batch (one captured graph)",
           "value": n * len(gr) / sum(gr), "unit": "configs/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": sum(gr) / len(gr) * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
           "data": "synthetic (seeded, generated on device)",
           "config": {"workload": "6x 16^3x32 + 6x 24^3x48 + 4x 32^3x64 periodic, eps U[0.2,0.4], "
                                  "APE levels {0,8,24}, R,T<=L/2", "parallelism": "1 GPU"},
           "sequential_configs_per_s": n * len(seq) / sum(seq),
           "graph_over_sequential": sum(seq) / sum(gr),
           "graph_equals_sequential_bitwise": bool(same)}
    for l in lats:
        l.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS) + ["c4"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default launch shapes (for profiling runs)")
    ap.add_argument("--batch", type=int, default=2,
                    help="configurations per captured graph in the batched leg (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=12,
                    help="requests per precision in the serving leg (0: skip)")
    ap.add_argument("--serve-sizes", default="12,16,24,32,48,64",
                    help="spatial extents L of the serving leg's L^3 x 2L lattices")
    ap.add_argument("--cpu-baseline-full", action="store_true",
                    help="only time whole 8^4 and 24^3x48 measurements of the CPU oracle per build "
                         "and thread count, print them as one JSON line and exit")
    args = ap.parse_args()
    if args.cpu_baseline_full:
        print(json.dumps(cpu_baseline_full(log)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); multi-GPU runs are not "
              "part of this build: independent configurations shard one per GPU with no "
              "communication.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "c4":
        batch_workload(args)
        return
    import wloop_amd as W
    cfg = WORKLOADS[args.workload]
    dims = cfg["dims"]
    parms = W.make_parms(cfg["nss"], cfg["T"], cfg["R"], alpha_ape=ALPHA_APE, proj_iter=PROJ_ITER,
                         alpha_action=ALPHA_ACTION)
    log(f"[bench] workload {args.workload}: {cfg['name']}")
    lat = W.Lattice(dims, cfg["bc"])
    lat.generate(cfg["seed"], cfg["eps"])
    lat.synchronize()
    log("[bench] field generated on device")

    from wloop_amd import autotune

    def describe(ch):
        d = {k: v for k, v in ch.items() if k not in ("trials", "cached")}
        d["kernel"] = W.KERNEL_NAMES[ch["kernel"]]
        return d

    tuned, choice = {}, {}
    for prec, name in ((W.FP64, "fp64"), (W.MIXED, "mixed")):
        if args.no_autotune:
            tuned[name] = "library default"
            continue
        choice[name] = autotune.tune(lat, parms, prec, cache=None, log=log)
        tuned[name] = describe(choice[name])
    # the plain FP64 path keeps the reference's structure: one contraction
    # launch per (R,T) over stored transporters (streaming or chain forms;
    # without autotuning the two-lane chain, the fastest of them at 48^3x96)
    if not args.no_autotune:
        # (on the FP64 path form 8 is the fused form)
        plain = min((t for t in choice["fp64"]["trials"]
                     if t["kernel"] not in (W.KERNEL_DEPTH, W.KERNEL_STORED9)),
                    key=lambda t: t["ms_per_rt"])
        tuned["fp64_plain"] = describe(plain)
    else:
        plain = {"kernel": W.KERNEL_CHAIN2, "wg": 0}
    autotune.apply(lat, plain, W.FP64)
    log(f"[bench] autotuned contraction: {tuned}")

    w64, t64, lat64, tim64, ls64 = time_path(lat, parms, W.FP64, args.steps, args.warmup)
    log(f"[bench] fp64 (plain): {t64 / args.steps * 1e3:.1f} ms/config  stages {tim64}")
    best64 = None
    if args.no_autotune or choice["fp64"]["kernel"] in (W.KERNEL_DEPTH, W.KERNEL_STORED9):
        # the fused time-slice form for FP64 (library default shape without autotuning)
        autotune.apply(lat, {"kernel": W.KERNEL_DEPTH, "wg": 0} if args.no_autotune else choice["fp64"],
                       W.FP64)
        w64b, t64b, lat64b, tim64b, ls64b = time_path(lat, parms, W.FP64, args.steps, args.warmup)
        log(f"[bench] fp64 (fused time slices): {t64b / args.steps * 1e3:.1f} ms/config  stages {tim64b}")
        best64 = {"value": args.steps / t64b, "ms_per_step": t64b / args.steps * 1e3,
                  "contraction": tuned["fp64"] if choice else "fused time slices, library default shape",
                  "latency_ms": {"p50": float(np.percentile(lat64b, 50)),
                                 "p99": float(np.percentile(lat64b, 99))},
                  "stage_ms_per_config": {k: v / args.steps for k, v in tim64b.items()},
                  "max_rel_err_vs_plain": rel_err(w64b, w64)}
    wmx, tmx, latmx, timmx, lsmx = time_path(lat, parms, W.MIXED, args.steps, args.warmup)
    log(f"[bench] mixed: {tmx / args.steps * 1e3:.1f} ms/config  stages {timmx}")
    err = rel_err(wmx, w64)
    log(f"[bench] mixed vs fp64 max rel err {err:.3e}")

    batched = None
    if args.batch > 0:
        # mixed path, args.batch configurations of the workload per replay of one
        # captured graph (fields regenerated on the device between replays,
        # outside the timed replays)
        lat.trim()
        lats = [lat] + [W.Lattice(dims, cfg["bc"]) for _ in range(args.batch - 1)]
        for l in lats[1:]:
            autotune.apply(l, autotune.current(lat, W.MIXED), W.MIXED)
        batch = W.Batch(lats, [parms] * args.batch, W.MIXED)
        bms = []
        errs = []
        for rep in range(args.warmup + args.steps):
            for i, l in enumerate(lats):
                l.generate(cfg["seed"] + 1000 * rep + i if i or rep else cfg["seed"], cfg["eps"])
            t0 = time.perf_counter()
            outs = batch.run()
            dt = time.perf_counter() - t0
            if rep >= args.warmup:
                bms.append(dt)
            if rep == 0:
                errs.append(rel_err(outs[0], w64))
        batch.close()
        for l in lats[1:]:
            l.close()
        tb = sum(bms)
        batched = {"configs_per_graph": args.batch, "value": args.batch * len(bms) / tb,
                   "ms_per_config": tb / (args.batch * len(bms)) * 1e3,
                   "latency_ms": {"p50": float(np.percentile(np.array(bms) * 1e3, 50)),
                                  "p99": float(np.percentile(np.array(bms) * 1e3, 99))},
                   "max_rel_err_vs_fp64": errs[0] if errs else None}
        log(f"[bench] batched mixed: {batched}")

    # drop-in boundary: there the gauge field arrives in host memory (openQCD's
    # udfld(), pageable): upload = host -> device copy + AoS -> SoA on the device
    u_host = lat.download()
    lat.upload(u_host)
    lat.synchronize()
    t0 = time.perf_counter()
    lat.upload(u_host)
    lat.synchronize()
    up_ms = (time.perf_counter() - t0) * 1e3
    del u_host
    host_input = {"upload_ms": up_ms, "bytes": int(np.prod(dims)) * 4 * 144,
                  "mixed_ms_per_config_incl_upload": tmx / args.steps * 1e3 + up_ms,
                  "mixed_configs_per_s_incl_upload": 1e3 / (tmx / args.steps * 1e3 + up_ms)}
    log(f"[bench] host input: {host_input}")

    V = int(np.prod(dims))
    prof_path = os.path.join(ROOT, "profiles", "kernel_profile_48c96_open.json")
    profile = {}
    if os.path.exists(prof_path) and args.workload == "c3":
        with open(prof_path) as f:
            profile = json.load(f)
    def starts(name):  # start slices of the fused launches of a path (library defaults: mixed 4, FP64 8)
        return (choice.get(name) or {}).get("starts") or (4 if name == "mixed" else 8)

    def form(name):  # contraction form of a path (library defaults: mixed 7, FP64 4)
        return (choice.get(name) or {}).get("kernel") or (W.KERNEL_STORED if name == "mixed" else W.KERNEL_DEPTH)

    # the committed profile measured the library's default launch shapes (its
    # command runs without autotuning): its traffic applies to a path that ran them
    def prof_for(name, prec):
        ch = choice.get(name)
        same = not ch or all(ch.get(k) == v for k, v in autotune.DEFAULT_SHAPE[prec].items())
        return profile if same else {}
    # (the plain FP64 path: every form that runs the two-lane chain on FP64)
    plain_prof = profile if plain["kernel"] in (W.KERNEL_CHAIN2, W.KERNEL_HYBRID, W.KERNEL_STORED) else {}
    rooflines = {"mixed": kernel_rooflines("mixed", parms, V, timmx, lsmx, args.steps, prof_for("mixed", W.MIXED),
                                           starts("mixed"), form("mixed")),
                 "fp64": kernel_rooflines("fp64", parms, V, tim64, ls64, args.steps, plain_prof, 8, plain["kernel"])}
    if best64:
        rooflines["fp64_best"] = kernel_rooflines("fp64", parms, V, tim64b, ls64b, args.steps,
                                                  prof_for("fp64", W.FP64), starts("fp64"), form("fp64"))
        best64["roofline"] = max(rooflines["fp64_best"], key=lambda e: e["ms_per_config"])
    primary = max(rooflines["mixed"], key=lambda e: e["ms_per_config"])

    out = {
        "metric": "configs/s at 48^3x96 (mixed vs FP64 GPU, rel err<1e-6); p50/p99 latency",
        "value": args.steps / tmx,
        "unit": "configs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tmx / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded Philox links exp(i eps H) with Haar gauge transform, generated on device)",
        "config": {"workload": cfg["name"], "lattice": list(dims), "bc": "open" if cfg["bc"] == 0 else "periodic",
                   "seed": cfg["seed"], "nss": cfg["nss"], "mwlt": cfg["T"], "mwlr": cfg["R"],
                   "proj_iter": PROJ_ITER, "alpha_action": list(ALPHA_ACTION), "alpha_ape": ALPHA_APE,
                   "parallelism": "1 GPU"},
        "latency_ms": {"p50": float(np.percentile(latmx, 50)), "p99": float(np.percentile(latmx, 99))},
        "fp64": {"value": args.steps / t64, "ms_per_step": t64 / args.steps * 1e3,
                 "latency_ms": {"p50": float(np.percentile(lat64, 50)), "p99": float(np.percentile(lat64, 99))},
                 "stage_ms_per_config": {k: v / args.steps for k, v in tim64.items()},
                 "roofline": max(rooflines["fp64"], key=lambda e: e["ms_per_config"])},
        "mixed_over_fp64": t64 / tmx,
        "fp64_best": best64,
        "mixed_over_fp64_best": (best64["ms_per_step"] / (tmx / args.steps * 1e3)) if best64 else t64 / tmx,
        "autotuned_contraction": tuned,
        "batched_mixed": batched,
        "mixed_max_rel_err_vs_fp64": err,
        "host_input": host_input,
        "stage_ms_per_config": {k: v / args.steps for k, v in timmx.items()},
        "roofline": primary,
        "roofline_by_kernel": rooflines,
    }
    if not args.no_cpu_baseline:
        po = None
        import oracle as O
        po = O.make_parms(cfg["nss"], cfg["T"], cfg["R"], alpha_ape=ALPHA_APE, proj_iter=PROJ_ITER)
        out["cpu_baseline"] = cpu_baseline(cfg, po)
        log(f"[bench] cpu baseline {out['cpu_baseline']}")
    else:
        out["cpu_baseline"] = None
    out["aux_kernels"] = aux_kernels(lat, 5)
    lat.close()
    sizes = tuple(int(x) for x in args.serve_sizes.split(","))
    out["serving"] = serving_leg(args.serve_requests, sizes) if args.serve_requests > 0 else None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
