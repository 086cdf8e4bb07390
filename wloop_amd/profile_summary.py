"""Summaries of rocprofv3 output for the Wilson-loop kernels.

    python -m wloop_amd.profile_summary KT_DIR FETCH_DIR WRITE_DIR SQ_DIRS OUT_JSON
    python -m wloop_amd.profile_summary --all KT_DIR FETCH_DIR WRITE_DIR

KT_DIR holds a ``--kernel-trace --stats`` run (``*_kernel_stats.csv``),
FETCH_DIR and WRITE_DIR ``--pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE`` runs of the
same command, SQ_DIRS a comma-separated list of ``--pmc`` runs of SQ counters
(``*_counter_collection.csv``).  The byte counters are calibrated per dispatch
against the FP64 -> FP32 copy kernel ``d2f`` (one complex number per thread:
16 B read and 8 B written per grid item).

Kernels are grouped by launch kind and precision ("ape_step/mixed",
"fused/mixed", "chain/fp64", "chain_build/mixed", ...; see kernel_key).  Prints
markdown tables (kernel time; HBM traffic per launch; SQ counters per launch
with the SIMD busy estimate) and writes OUT_JSON with, per key, the
calibrated HBM bytes per launch, the average launch time and the SQ counters
per launch -- the figures bench.py reads for the ``traffic`` field of its
roofline entries.  With ``--all`` it prints the calibrated traffic and rate of
every kernel of unfiltered counter runs.

The clock the chip ran at is SQ_BUSY_CYCLES / (32 shader engines x launch
time).  SQ_INSTS_VALU counts the matrix-core instructions too.  SIMD busy
estimate = (VALU cycles + SQ_VALU_MFMA_BUSY_CYCLES) / (1024 SIMDs x cycles of
the launch), with the VALU cycles of a wave64 instruction as measured on gfx950
by wloop_amd/issue_rates (16 waves on one CU, independent chains): 4.9 for FP64
arithmetic (SQ_INSTS_VALU_FMA/MUL/ADD_F64) and conversions (SQ_INSTS_VALU_CVT),
VALU_OTHER_CYCLES = 2.5 for the rest (FP32 and integer: 2.25-2.8 in the 4-byte
encodings, with and without VGPR bank conflicts, 2.95-3.1 in the 8-byte ones;
DPP moves 4.9 are counted at 2.5).  Matrix-core and VALU instructions of one
SIMD do not overlap on this chip (issue_rates: 8 v_mfma_f32_4x4x1f32 and 16
v_fmac_f32 take longer than the two apart), so the two are added.
"""
import collections
import csv
import glob
import json
import os
import re
import sys

SIMDS = 1024
SHADER_ENGINES = 32
VALU_F64_CYCLES = 4.9    # FP64 arithmetic and conversions, per wave64 instruction
VALU_OTHER_CYCLES = 2.5  # FP32 and integer


def _csv(d, suffix):
    files = glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True)
    if not files:
        raise FileNotFoundError(f"no *{suffix} under {d}")
    with open(files[0]) as f:
        return list(csv.DictReader(f))


def _targs(name):
    if "<" not in name:
        return []
    return [a.strip() for a in name[name.index("<") + 1:name.index(">")].split(",")]


def kernel_key(name):
    """'<kind>/<precision>' of a kernel symbol, else None.  Precision 'mixed'
    for FP32 link storage, 'fp64' for FP64; the fused launches of the mixed
    path that carry FP64 open loops (T >= 9) are 'fused64/mixed'; the chain
    launch that also forms the spatial transporters is 'chain_build'."""
    a = _targs(name)
    if re.search(r"\bape_step<", name):
        return "ape_step/" + ("mixed" if a[1] == "float" else "fp64")
    m = re.search(r"\bhyp_level(\d)<", name)
    if m:
        return f"hyp{m.group(1)}/" + ("mixed" if a[0] == "float" else "fp64")
    if re.search(r"\bcontract_depth<", name):
        ts, tc = a[2], a[3]
        if len(a) > 4 and a[4] == "true":
            return "stored/mixed"
        if ts == "double":
            return "fused/fp64"
        return "fused/mixed" if tc == "float" else "fused64/mixed"
    # template position of the link storage type of each per-(R,T) form
    for stem, its in (("contract_chain1", 1), ("contract_chain2", 1), ("contract_stream", 1),
                      ("contract_chain", 3)):
        if re.search(r"\b" + stem + r"<", name):
            prec = "fp64" if a[its] == "double" else "mixed"
            build = a[-1] == "true" and stem in ("contract_chain1", "contract_chain2")
            return ("chain_build/" if build else "chain/") + prec
    if re.search(r"\bspace_transporter<", name):
        return "space_tr/" + ("mixed" if a[0] == "float" else "fp64")
    if re.search(r"\btime_transporter<", name):
        return "time_tr/" + ("mixed" if a[0] == "float" else "fp64")
    if re.search(r"\bfinish_wloop\b", name):
        return "finish/any"
    return None


def short(name, n=64):
    name = re.sub(r"\(.*", "", name)
    return name if len(name) <= n else name[:n - 1] + "~"


def stats_table(kt_dir, top=16):
    rows = _csv(kt_dir, "_kernel_stats.csv")
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = ["| kernel | calls | avg ms | total ms | % |", "|---|---:|---:|---:|---:|"]
    for r in rows[:top]:
        out.append(f"| `{short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e6:.3f} | "
                   f"{float(r['TotalDurationNs']) / 1e6:.1f} | {float(r['Percentage']):.1f} |")
    return "\n".join(out), rows


def avg_ms_by_key(kt_dir):
    agg = collections.defaultdict(lambda: [0.0, 0])
    for r in _csv(kt_dir, "_kernel_stats.csv"):
        k = kernel_key(r["Name"])
        if k:
            agg[k][0] += float(r["TotalDurationNs"]) / 1e6
            agg[k][1] += int(r["Calls"])
    return {k: t / n for k, (t, n) in agg.items() if n}


def _per_dispatch(d, counter):
    rows = _csv(d, "_counter_collection.csv")
    out = []
    for r in rows:
        if r["Counter_Name"] == counter:
            out.append((r["Kernel_Name"], int(r["Grid_Size"]), float(r["Counter_Value"]) * 1024.0))
    return out


def _calibration(fetch, write):
    cf = [16.0 * g / b for n, g, b in fetch if re.search(r"\bd2f\b", n) and b > 0]
    cw = [8.0 * g / b for n, g, b in write if re.search(r"\bd2f\b", n) and b > 0]
    if not cf or not cw:
        raise RuntimeError("no d2f dispatch to calibrate against")
    return sum(cf) / len(cf), sum(cw) / len(cw)


def traffic(fetch_dir, write_dir):
    fetch = _per_dispatch(fetch_dir, "FETCH_SIZE")
    write = _per_dispatch(write_dir, "WRITE_SIZE")
    kf, kw = _calibration(fetch, write)
    agg = collections.defaultdict(lambda: [[], []])
    names = {}
    for idx, rows in ((0, fetch), (1, write)):
        for n, g, b in rows:
            key = kernel_key(n)
            if key:
                agg[key][idx].append(b * (kf if idx == 0 else kw))
                names[key] = short(n)
    per_launch = {k: (sum(f) / len(f) if f else 0.0) + (sum(w) / len(w) if w else 0.0)
                  for k, (f, w) in agg.items()}
    table = ["| kernel (key) | launches | fetched GB/launch | written GB/launch |", "|---|---:|---:|---:|"]
    for k, (f, w) in sorted(agg.items()):
        table.append(f"| `{names[k]}` ({k}) | {len(f)} | {sum(f) / max(len(f), 1) / 1e9:.3f} | "
                     f"{sum(w) / max(len(w), 1) / 1e9:.3f} |")
    return kf, kw, per_launch, "\n".join(table)


def sq_counters(dirs):
    """{key: {counter: mean value per launch}} over the SQ counter runs"""
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    for d in dirs:
        for r in _csv(d, "_counter_collection.csv"):
            k = kernel_key(r["Kernel_Name"])
            if k:
                agg[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in agg.items()}


def valu_cycles(c):
    """SIMD cycles of a launch's VALU instructions (matrix-core ones excluded)
    under the measured issue costs: FP64 arithmetic and conversions at
    VALU_F64_CYCLES, the rest at VALU_OTHER_CYCLES"""
    f64 = sum(c.get(x, 0.0) for x in ("SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_ADD_F64",
                                       "SQ_INSTS_VALU_CVT"))
    rest = max(c.get("SQ_INSTS_VALU", 0.0) - c.get("SQ_INSTS_MFMA", 0.0) - f64, 0.0)
    return VALU_F64_CYCLES * f64 + VALU_OTHER_CYCLES * rest


def sq_table(sq, avg_ms):
    cols = ["SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_VALU_FMA_F32", "SQ_INSTS_VALU_FMA_F64",
            "SQ_INSTS_VALU_MUL_F32", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_CVT",
            "SQ_INSTS_VMEM_RD", "SQ_INSTS_LDS", "SQ_INSTS_MFMA"]
    out = ["| key | avg ms | " + " | ".join(c.replace("SQ_INSTS_", "").replace("SQ_", "") for c in cols) +
           " | FMA share | clock GHz | VALU busy (model) | MFMA busy | SIMD busy (model) |",
           "|---|" + "---:|" * (len(cols) + 6)]
    derived = {}
    for k in sorted(sq):
        c = sq[k]
        t = avg_ms.get(k)
        valu = c.get("SQ_INSTS_VALU", 0.0) - c.get("SQ_INSTS_MFMA", 0.0)
        fma = c.get("SQ_INSTS_VALU_FMA_F32", 0.0) + c.get("SQ_INSTS_VALU_FMA_F64", 0.0)
        busy = c.get("SQ_BUSY_CYCLES")
        cyc = busy / SHADER_ENGINES if busy else None  # cycles of the launch at the running clock
        clock = cyc / (t * 1e-3) / 1e9 if (cyc and t) else None
        vb = valu_cycles(c) / (SIMDS * cyc) if cyc else None
        mb = c.get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0) / (SIMDS * cyc) if cyc else None
        derived[k] = {"fma_share": fma / valu if valu else None, "clock_ghz": clock, "valu_busy": vb,
                      "mfma_busy": mb, "simd_busy": (vb + mb) if vb is not None else None}

        def f(v):
            return f"{v:.2f}" if v is not None else "--"
        cells = [f"{c.get(x, 0.0):.4g}" for x in cols]
        d = derived[k]
        out.append(f"| {k} | {t:.3f} | " + " | ".join(cells) +
                   f" | {f(d['fma_share'])} | {f(clock)} | {f(vb)} | {f(mb)} | {f(d['simd_busy'])} |"
                   if t else f"| {k} | -- | " + " | ".join(cells) + " | -- | -- | -- | -- | -- |")
    return "\n".join(out), derived


def all_traffic(kt_dir, fetch_dir, write_dir):
    """markdown table of calibrated HBM bytes and rates per launch for every
    kernel of FETCH_DIR / WRITE_DIR runs made without a kernel filter"""
    fetch = _per_dispatch(fetch_dir, "FETCH_SIZE")
    write = _per_dispatch(write_dir, "WRITE_SIZE")
    kf, kw = _calibration(fetch, write)
    avg = {short(r["Name"], 200): float(r["AverageNs"]) / 1e6 for r in _csv(kt_dir, "_kernel_stats.csv")}
    agg = collections.defaultdict(lambda: [[], []])
    for idx, rows in ((0, fetch), (1, write)):
        for n, g, b in rows:
            agg[short(n, 200)][idx].append(b * (kf if idx == 0 else kw))
    out = [f"Calibration against `d2f`: FETCH_SIZE x {kf:.4f}, WRITE_SIZE x {kw:.4f}.", "",
           "| kernel | launches | fetched GB/launch | written GB/launch | avg ms | TB/s |",
           "|---|---:|---:|---:|---:|---:|"]
    for k, (f, w) in sorted(agg.items(), key=lambda kv: -sum(kv[1][0])):
        fb = sum(f) / len(f) if f else 0.0
        wb = sum(w) / len(w) if w else 0.0
        t = avg.get(k, 0.0)
        rate = f"{(fb + wb) / t / 1e9:.2f}" if t else "--"
        out.append(f"| `{short(k)}` | {len(f)} | {fb / 1e9:.2f} | {wb / 1e9:.2f} | {t:.3f} | {rate} |")
    return "\n".join(out)


def main(argv):
    if len(argv) == 5 and argv[1] == "--all":
        print(all_traffic(*argv[2:]))
        return 0
    if len(argv) != 6:
        print(__doc__, file=sys.stderr)
        return 2
    kt_dir, fetch_dir, write_dir, sq_dirs, out_json = argv[1:]
    st, _ = stats_table(kt_dir)
    avg_ms = avg_ms_by_key(kt_dir)
    kf, kw, per_launch, tt = traffic(fetch_dir, write_dir)
    sq = sq_counters([d for d in sq_dirs.split(",") if d])
    sqt, derived = sq_table(sq, avg_ms)
    print("## Kernel time\n\n" + st + "\n")
    print(f"## HBM traffic per launch\n\nCalibration against `d2f`: FETCH_SIZE x {kf:.4f}, "
          f"WRITE_SIZE x {kw:.4f}.\n\n" + tt + "\n")
    print("## SQ counters per launch\n\n" + sqt + "\n")
    with open(out_json, "w") as f:
        json.dump({"fetch_calibration": kf, "write_calibration": kw, "avg_launch_ms": avg_ms,
                   "bytes_per_launch": per_launch, "sq_per_launch": sq, "sq_derived": derived},
                  f, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
