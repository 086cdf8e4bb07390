"""Launch-shape autotuner for the Wilson-loop contraction.

For a lattice shape, boundary condition, number of operator levels, T range and
precision, it times every candidate (contraction form x workgroup parameter,
and for the fused time-slice form x time-slice depth x start slices per
workgroup) on the resident field with HIP events, the fastest of three runs,
and keeps the one with the least time per (R,T) pair of the stages the form
decides: the contraction launches and the spatial transporters (formed by a
separate pass, or inside the first chain launch of an R: both happen once per
R >= 2, in the trial as in the full run) per (R,T) pair of the trial, plus the
time transporters (one chain per measurement, more of it stored for the
per-(R,T) forms) spread over the (R,T) pairs of the full run.  Results are
cached in a JSON file keyed by those properties, so a shape is tuned once per
machine.  The choice changes only the launch shape: results agree to
floating-point rounding and summation order (checked in
tests/test_gpu_tuning.py).
"""
import json
import os

import wloop_amd as W


def _cands():
    out = [{"kernel": W.KERNEL_STREAM, "wg": wg} for wg in (4, 8, 16, 32)]
    out += [{"kernel": W.KERNEL_CHAIN, "wg": wg} for wg in (512, 2048)]
    out += [{"kernel": W.KERNEL_CHAIN1, "wg": wg} for wg in (256, 512, 1024)]
    out += [{"kernel": W.KERNEL_CHAIN2, "wg": wg} for wg in (256, 512, 1024)]
    out += [{"kernel": W.KERNEL_DEPTH, "wg": wg, "depth": d, "starts": st}
            for d in (8, 16) for st in (4, 8) for wg in (32, 64, 128)]
    out += [{"kernel": kern, "wg": wg, "depth": 8, "starts": st}
            for kern in (W.KERNEL_HYBRID, W.KERNEL_STORED, W.KERNEL_STORED9) for st in (4, 8)
            for wg in (256, 512)]
    return out


CANDIDATES = _cands()


def path_candidates(precision):
    """the candidates that are distinct launch shapes on a path: on the FP64
    path forms 6 and 7 run the two-lane chain (their FP32 blocks exist on the
    mixed path only), so they repeat candidates of form 5 there"""
    if precision == W.MIXED:
        return CANDIDATES
    return [c for c in CANDIDATES if c["kernel"] not in (W.KERNEL_HYBRID, W.KERNEL_STORED)]
# the library's default launch shape per precision (contract_choice in
# device/api.hip); a trial candidate replaces it only when it is faster by more
# than KEEP_DEFAULT_WITHIN, about the spread of single trial runs, so that the
# choice does not follow the noise between shapes of equal speed
DEFAULT_SHAPE = {W.MIXED: {"kernel": W.KERNEL_STORED, "wg": 512, "depth": 8, "starts": 8},
                 W.FP64: {"kernel": W.KERNEL_DEPTH, "wg": 64, "depth": 16, "starts": 8}}
KEEP_DEFAULT_WITHIN = 0.02
DEFAULT_CACHE = os.environ.get(
    "WLOOP_AMD_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "wloop_amd", "tune.json"))


def _key(lat, parms, precision):
    return "{}x{}x{}x{}-bc{}-msl{}-T{}-R{}-{}".format(*lat.dims, lat.bc, parms.msl, parms.mwlt, parms.mwlr,
                                                      "mixed" if precision == W.MIXED else "fp64")


def _params(precision):
    if precision == W.MIXED:
        return (W.TUNE_MIXED_CONTRACT_KERNEL, W.TUNE_MIXED_CONTRACT_WG, W.TUNE_MIXED_CONTRACT_DEPTH,
                W.TUNE_MIXED_CONTRACT_STARTS)
    return (W.TUNE_FP64_CONTRACT_KERNEL, W.TUNE_FP64_CONTRACT_WG, W.TUNE_FP64_CONTRACT_DEPTH,
            W.TUNE_FP64_CONTRACT_STARTS)


def _load(cache):
    try:
        with open(cache) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def apply(lat, choice, precision):
    pk, pw, pd, ps = _params(precision)
    lat.tune(pk, choice["kernel"])
    lat.tune(pw, choice["wg"])
    lat.tune(pd, choice.get("depth", 0))
    lat.tune(ps, choice.get("starts", 0))


def current(lat, precision):
    """the launch shape a context uses now (0 = library default)"""
    pk, pw, pd, ps = _params(precision)
    return {"kernel": lat.tune_get(pk), "wg": lat.tune_get(pw), "depth": lat.tune_get(pd),
            "starts": lat.tune_get(ps)}


def tune(lat, parms, precision, cache=DEFAULT_CACHE, candidates=None, force=False, log=None):
    """Pick and apply the fastest contraction launch shape; returns the choice
    {"kernel", "wg"[, "depth", "starts"], "ms_per_rt", "trials"}, ms_per_rt
    being the projected time per (R,T) pair of the stages the form decides."""
    key = _key(lat, parms, precision)
    db = _load(cache) if cache else {}
    if not force and key in db:
        apply(lat, db[key], precision)
        return dict(db[key], cached=True)
    # a short measurement with the real level count and T range: distinct
    # levels one APE step apart, two values of R
    trial = W.make_parms([0] + [1] * (parms.msl - 1), parms.mwlt, min(parms.mwlr, 2),
                         proj_iter=parms.proj_iter)
    cands = candidates or path_candidates(precision)
    trials = []
    for cand in cands:
        apply(lat, cand, precision)
        lat.wloop(trial, precision)  # warm-up and allocation
        # the fastest of three runs (the forms differ by a few per cent, about
        # the spread of single runs)
        per_rt = None
        for _ in range(3):
            lat.timers_reset()
            lat.wloop(trial, precision)
            st = lat.timers()
            v = ((st["contract"] + st["space_transporter"]) / (trial.mwlt * trial.mwlr) +
                 st["time_transporter"] / (parms.mwlt * parms.mwlr))
            per_rt = v if per_rt is None else min(per_rt, v)
        trials.append(dict(cand, ms_per_rt=per_rt))
        if log:
            log(f"[autotune] {key} {cand}: {per_rt:.3f} ms per (R,T)")
    best = min(trials, key=lambda t: t["ms_per_rt"])
    dflt = DEFAULT_SHAPE.get(precision)
    for t in trials:
        if dflt and all(t.get(k) == v for k, v in dflt.items()) and \
                t["ms_per_rt"] <= best["ms_per_rt"] * (1.0 + KEEP_DEFAULT_WITHIN):
            best = t
    choice = dict(best, trials=trials)
    apply(lat, choice, precision)
    if cache:
        db[key] = choice
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        tmp = cache + ".tmp"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1)
        os.replace(tmp, cache)
    return dict(choice, cached=False)
