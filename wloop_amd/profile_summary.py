"""Summaries of rocprofv3 output for the Wilson-loop kernels.

    python -m wloop_amd.profile_summary KT_DIR FETCH_DIR WRITE_DIR OUT_JSON
    python -m wloop_amd.profile_summary --all KT_DIR FETCH_DIR WRITE_DIR

KT_DIR holds a ``--kernel-trace --stats`` run (``*_kernel_stats.csv``),
FETCH_DIR and WRITE_DIR ``--pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE`` runs of the
same command (``*_counter_collection.csv``).  The byte counters are calibrated
per dispatch against the FP64 -> FP32 copy kernel ``d2f`` (one complex number
per thread: 16 B read and 8 B written per grid item).  Prints markdown tables
(kernel times; HBM traffic per launch of the contraction kernels) and writes
OUT_JSON with the calibrated bytes per launch keyed "<form>/<precision>", the
keys bench.py reads for the ``roofline.traffic`` field.  With ``--all`` it prints
the calibrated traffic and rate of every kernel of unfiltered counter runs.
"""
import collections
import csv
import glob
import json
import os
import re
import sys

FORMS = (("contract_stream", 1), ("contract_chain1", 3), ("contract_chain2", 5), ("contract_chain", 2),
         ("contract_depth", 4))


def _csv(d, suffix):
    files = glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True)
    if not files:
        raise FileNotFoundError(f"no *{suffix} under {d}")
    with open(files[0]) as f:
        return list(csv.DictReader(f))


def kernel_key(name):
    """'<form>/<precision>' of a contraction kernel symbol, else None; the
    one-lane chain launch that also forms the spatial transporters is form '3b'"""
    for stem, form in FORMS:
        if re.search(r"\b" + stem + r"<", name):
            args = [a.strip() for a in name[name.index("<") + 1:name.index(">")].split(",")]
            prec = "fp64" if args[-2] == "double" else "mixed"
            if stem == "contract_chain1" and args[-1] == "true":
                form = "3b"
            return f"{form}/{prec}"
    return None


def short(name, n=64):
    name = re.sub(r"\(.*", "", name)
    return name if len(name) <= n else name[:n - 1] + "~"


def stats_table(kt_dir, top=14):
    rows = _csv(kt_dir, "_kernel_stats.csv")
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = ["| kernel | calls | avg ms | total ms | % |", "|---|---:|---:|---:|---:|"]
    for r in rows[:top]:
        out.append(f"| `{short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e6:.3f} | "
                   f"{float(r['TotalDurationNs']) / 1e6:.1f} | {float(r['Percentage']):.1f} |")
    return "\n".join(out), rows


def _per_dispatch(d, counter):
    rows = _csv(d, "_counter_collection.csv")
    out = []
    for r in rows:
        if r["Counter_Name"] == counter:
            out.append((r["Kernel_Name"], int(r["Grid_Size"]), float(r["Counter_Value"]) * 1024.0))
    return out


def traffic(fetch_dir, write_dir):
    fetch = _per_dispatch(fetch_dir, "FETCH_SIZE")
    write = _per_dispatch(write_dir, "WRITE_SIZE")
    cf = [16.0 * g / b for n, g, b in fetch if re.search(r"\bd2f\b", n) and b > 0]
    cw = [8.0 * g / b for n, g, b in write if re.search(r"\bd2f\b", n) and b > 0]
    if not cf or not cw:
        raise RuntimeError("no d2f dispatch to calibrate against")
    kf, kw = sum(cf) / len(cf), sum(cw) / len(cw)
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
    table = ["| kernel | launches | fetched GB/launch | written MB/launch |", "|---|---:|---:|---:|"]
    for k, (f, w) in sorted(agg.items()):
        table.append(f"| `{names[k]}` ({k}) | {len(f)} | {sum(f) / len(f) / 1e9:.2f} | "
                     f"{sum(w) / max(len(w), 1) / 1e6:.2f} |")
    return kf, kw, per_launch, "\n".join(table)


def all_traffic(kt_dir, fetch_dir, write_dir):
    """markdown table of calibrated HBM bytes and rates per launch for every
    kernel of FETCH_DIR / WRITE_DIR runs made without a kernel filter"""
    fetch = _per_dispatch(fetch_dir, "FETCH_SIZE")
    write = _per_dispatch(write_dir, "WRITE_SIZE")
    kf = [16.0 * g / b for n, g, b in fetch if re.search(r"\bd2f\b", n) and b > 0]
    kw = [8.0 * g / b for n, g, b in write if re.search(r"\bd2f\b", n) and b > 0]
    if not kf or not kw:
        raise RuntimeError("no d2f dispatch to calibrate against")
    kf, kw = sum(kf) / len(kf), sum(kw) / len(kw)
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
    if len(argv) != 5:
        print(__doc__, file=sys.stderr)
        return 2
    kt_dir, fetch_dir, write_dir, out_json = argv[1:]
    st, _ = stats_table(kt_dir)
    kf, kw, per_launch, tt = traffic(fetch_dir, write_dir)
    print("## Kernel time\n\n" + st + "\n")
    print(f"## HBM traffic per launch\n\nCalibration against `d2f`: FETCH_SIZE x {kf:.4f}, "
          f"WRITE_SIZE x {kw:.4f}.\n\n" + tt + "\n")
    with open(out_json, "w") as f:
        json.dump({"fetch_calibration": kf, "write_calibration": kw, "bytes_per_launch": per_launch},
                  f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
