"""Launch-shape autotuner for the Wilson-loop contraction.

For a lattice shape, boundary condition, number of operator levels and
precision, it times every candidate (contraction form x workgroup parameter)
on the resident field with HIP events and keeps the fastest; results are cached
in a JSON file keyed by those properties, so a shape is tuned once per machine.
The choice changes only the launch shape: results agree to floating-point
summation order (checked in tests/test_gpu_tuning.py).
"""
import json
import os

import wloop_amd as W

CANDIDATES = {
    W.KERNEL_STREAM: (4, 8, 16, 32),
    W.KERNEL_CHAIN: (512, 2048),
    W.KERNEL_CHAIN1: (256, 512, 1024),
}
DEFAULT_CACHE = os.environ.get(
    "WLOOP_AMD_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "wloop_amd", "tune.json"))


def _key(lat, parms, precision):
    return "{}x{}x{}x{}-bc{}-msl{}-{}".format(*lat.dims, lat.bc, parms.msl,
                                             "mixed" if precision == W.MIXED else "fp64")


def _params(precision):
    if precision == W.MIXED:
        return W.TUNE_MIXED_CONTRACT_KERNEL, W.TUNE_MIXED_CONTRACT_WG
    return W.TUNE_FP64_CONTRACT_KERNEL, W.TUNE_FP64_CONTRACT_WG


def _load(cache):
    try:
        with open(cache) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def apply(lat, choice, precision):
    pk, pw = _params(precision)
    lat.tune(pk, choice["kernel"])
    lat.tune(pw, choice["wg"])


def tune(lat, parms, precision, cache=DEFAULT_CACHE, candidates=None, force=False, log=None):
    """Pick and apply the fastest contraction launch shape; returns the choice
    {"kernel", "wg", "ms_per_launch", "trials"}."""
    key = _key(lat, parms, precision)
    db = _load(cache) if cache else {}
    if not force and key in db:
        apply(lat, db[key], precision)
        return dict(db[key], cached=True)
    # a short measurement with the real level count: no APE steps, a few R,T
    trial = W.make_parms([0] * parms.msl, min(parms.mwlt, 4), min(parms.mwlr, 4),
                         proj_iter=parms.proj_iter)
    cands = candidates or CANDIDATES
    trials = []
    for kernel, wgs in cands.items():
        for wg in wgs:
            apply(lat, {"kernel": kernel, "wg": wg}, precision)
            lat.wloop(trial, precision)  # warm-up and allocation
            lat.timers_reset()
            for _ in range(2):
                lat.wloop(trial, precision)
            n, ms = lat.contract_stats()
            trials.append({"kernel": kernel, "wg": wg, "ms_per_launch": ms / max(n, 1)})
            if log:
                log(f"[autotune] {key} kernel={kernel} wg={wg}: {ms / max(n, 1):.3f} ms/launch")
    best = min(trials, key=lambda t: t["ms_per_launch"])
    choice = {"kernel": best["kernel"], "wg": best["wg"], "ms_per_launch": best["ms_per_launch"],
              "trials": trials}
    apply(lat, choice, precision)
    if cache:
        db[key] = choice
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        tmp = cache + ".tmp"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1)
        os.replace(tmp, cache)
    return dict(choice, cached=False)
