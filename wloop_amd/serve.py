"""Measurement daemon: serves a stream of configurations of uneven lattice sizes.

Each request names a lattice shape, boundary condition, generator seed and eps;
the daemon generates the configuration on the GPU and runs the whole
measurement (HYP, operator levels, transporters, W_ij(R,T;x0)) there.

* State stays on the device: one context (gauge field, scratch arena, partial
  sums) per (shape, bc, precision) is kept after a request and reused by the
  next request of that shape.
* Admission by HBM footprint: a request starts only when the device bytes its
  context needs (wlg_footprint) fit in the budget next to the contexts that
  already exist; idle contexts are evicted least-recently-used first.  A shape
  whose default layout does not fit the budget even alone gets a context with
  the lean layout (wlg_footprint_lean: fused contraction, time transporters
  only at block starts), and only a request that fits neither is rejected.
* Workers run requests concurrently, each context on its own HIP stream; the
  next request is the one of least predicted cost (with an ageing credit and
  a bonus for a resident context of its shape), or the oldest with --policy fifo.
* Reported: p50/p99 latency (arrival to W on the host), configs/s, service
  times, rejections.

    python -m wloop_amd.serve --rate 4 --requests 200 --precision mixed
"""
import argparse
import collections
import json
import random
import sys
import threading
import time

import numpy as np

import wloop_amd as W

SIZES = (12, 16, 24, 32, 48, 64)
NSS = (0, 8, 16, 24)


class Request:
    __slots__ = ("rid", "dims", "bc", "seed", "eps", "parms", "precision", "t_arrival", "t_start",
                 "t_done", "result", "error")

    def __init__(self, rid, dims, bc, seed, eps, parms, precision):
        self.rid, self.dims, self.bc, self.seed, self.eps = rid, tuple(dims), bc, seed, eps
        self.parms, self.precision = parms, precision
        self.t_arrival = self.t_start = self.t_done = None
        self.result = self.error = None


def standard_request(rid, L, bc, precision, seed=None, eps=0.3, nss=NSS):
    """A request of the serving workload: L^3 x 2L lattice, APE levels nss, R,T up to L/2."""
    h = L // 2
    p = W.make_parms(list(nss), h, h, alpha_ape=0.5, proj_iter=4, alpha_action=(1.0, 1.0, 0.5))
    return Request(rid, (2 * L, L, L, L), bc, rid if seed is None else seed, eps, p, precision)


def context_layout(req, budget):
    """(lean, bytes) of the context a request needs: the default layout if it
    fits the budget by itself, else the lean layout (which may still not fit:
    admission then rejects the request)"""
    nbytes = W.footprint(req.dims, req.parms, req.precision)
    if nbytes <= budget:
        return False, nbytes
    return True, W.footprint(req.dims, req.parms, req.precision, lean=True)


class Admission:
    """HBM accounting of device contexts: acquire() returns an existing idle
    context of the key, or the right to create one once its bytes fit next to
    the existing contexts (evicting idle ones, least recently used first)."""

    def __init__(self, budget_bytes):
        self.budget = budget_bytes
        self.used = 0
        self.idle = collections.OrderedDict()   # handle -> (key, nbytes), LRU order
        self.busy = {}                          # handle -> (key, nbytes)
        self.cv = threading.Condition()
        self.evicted = 0

    def acquire(self, key, nbytes, close):
        """-> (handle or None, must_create: bool); raises MemoryError if the
        request can never fit.  close(handle) releases an evicted context."""
        with self.cv:
            while True:
                for h, (k, nb) in self.idle.items():
                    if k == key:
                        del self.idle[h]
                        self.busy[h] = (k, nb)
                        return h, False
                if nbytes > self.budget:
                    raise MemoryError(f"needs {nbytes / 1e9:.1f} GB, budget {self.budget / 1e9:.1f} GB")
                while self.used + nbytes > self.budget and self.idle:
                    h, (k, nb) = self.idle.popitem(last=False)
                    self.used -= nb
                    self.evicted += 1
                    close(h)
                if self.used + nbytes <= self.budget:
                    self.used += nbytes
                    return None, True
                self.cv.wait()

    def has_idle(self, key):
        with self.cv:
            return any(k == key for k, _ in self.idle.values())

    def register(self, handle, key, nbytes):
        with self.cv:
            self.busy[handle] = (key, nbytes)

    def release(self, handle):
        with self.cv:
            self.idle[handle] = self.busy.pop(handle)
            self.cv.notify_all()

    def abandon(self, nbytes):
        """a context that was admitted could not be created"""
        with self.cv:
            self.used -= nbytes
            self.cv.notify_all()


def expected_ms(req):
    """Predicted service time of a request on one MI355X (ms): per-site costs of
    HYP and the transporters, of an APE step and of one (R,T) contraction,
    measured at 48^3x96 (mixed path: 4.0 + 0.25 per APE step + 0.19 per (R,T)
    ns per site; the FP64 path takes about 1.4 times as long), plus a fixed
    launch overhead."""
    V = float(np.prod(req.dims))
    p = req.parms
    steps = sum(p.nss[:p.msl])
    ns_per_site = 4.0 + 0.25 * steps + 0.19 * p.mwlr * p.mwlt
    return 20.0 + V * ns_per_site * 1e-6 * (1.0 if req.precision == W.MIXED else 1.4)


class Scheduler:
    """Pending requests.  "fifo" serves them in arrival order; "sjf" serves the
    request with the least predicted cost: expected_ms, plus the cost of
    creating its context when no idle context of its shape is resident, minus
    an ageing credit of `age_weight` ms per ms waited (no request waits
    forever)."""

    def __init__(self, policy, resident, age_weight=0.05, create_ms_per_gb=2.0):
        self.policy, self.resident = policy, resident
        self.age_weight, self.create_ms_per_gb = age_weight, create_ms_per_gb
        self.pending = []
        self.cv = threading.Condition()
        self.open_items = 0

    def put(self, req):
        with self.cv:
            self.pending.append(req)
            self.open_items += 1
            self.cv.notify()

    def _score(self, req, now):
        cost = expected_ms(req)
        if not self.resident((req.dims, req.bc, req.precision)):
            cost += 20.0 + self.create_ms_per_gb * W.footprint(req.dims, req.parms, req.precision) / 1e9
        return cost - self.age_weight * (now - req.t_arrival) * 1e3

    def get(self):
        with self.cv:
            while not self.pending:
                self.cv.wait()
            if self.policy == "fifo" or self.pending[0] is None:
                return self.pending.pop(0)
            now = time.perf_counter()
            live = [r for r in self.pending if r is not None]
            best = min(live, key=lambda r: self._score(r, now))
            self.pending.remove(best)
            return best

    def stop_one(self):
        with self.cv:
            self.pending.append(None)
            self.cv.notify()

    def task_done(self):
        with self.cv:
            self.open_items -= 1
            self.cv.notify_all()

    def join(self):
        with self.cv:
            while self.open_items > 0:
                self.cv.wait()


class Server:
    def __init__(self, budget_bytes, workers=2, device=0, policy="sjf"):
        self.adm = Admission(budget_bytes)
        self.device = device
        self.q = Scheduler(policy, self.adm.has_idle)
        self.done = []
        self.lock = threading.Lock()
        self.threads = [threading.Thread(target=self._worker, daemon=True) for _ in range(workers)]
        for t in self.threads:
            t.start()

    def submit(self, req):
        req.t_arrival = time.perf_counter()
        self.q.put(req)

    def _worker(self):
        while True:
            req = self.q.get()
            if req is None:
                return
            try:
                self._serve(req)
            except Exception as e:  # noqa: BLE001 -- reported per request
                req.error = f"{type(e).__name__}: {e}"
            req.t_done = time.perf_counter()
            with self.lock:
                self.done.append(req)
            self.q.task_done()

    def _serve(self, req):
        key = (req.dims, req.bc, req.precision)
        lean, nbytes = context_layout(req, self.adm.budget)
        lat, create = self.adm.acquire(key, nbytes, lambda h: h.close())
        if create:
            try:
                lat = W.Lattice(req.dims, req.bc, self.device)
                if lean:
                    lat.tune(W.TUNE_LEAN, 1)
            except Exception:
                self.adm.abandon(nbytes)
                raise
            self.adm.register(lat, key, nbytes)
        try:
            req.t_start = time.perf_counter()
            lat.generate(req.seed, req.eps)
            req.result = lat.measure(req.parms, req.precision)
        finally:
            self.adm.release(lat)

    def drain(self):
        self.q.join()

    def stop(self):
        for _ in self.threads:
            self.q.stop_one()
        for t in self.threads:
            t.join()
        for h in list(self.adm.idle):
            h.close()


def summarize(reqs, t0):
    ok = [r for r in reqs if r.error is None]
    lat = np.array([(r.t_done - r.t_arrival) * 1e3 for r in ok]) if ok else np.zeros(1)
    svc = np.array([(r.t_done - r.t_start) * 1e3 for r in ok]) if ok else np.zeros(1)
    span = max(r.t_done for r in reqs) - t0 if reqs else 1.0
    by_size = collections.defaultdict(list)
    for r in ok:
        by_size["x".join(map(str, r.dims[::-1]))].append((r.t_done - r.t_start) * 1e3)
    return {"requests": len(reqs), "completed": len(ok), "rejected": len(reqs) - len(ok),
            "configs_per_s": len(ok) / span,
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
            "service_ms": {"p50": float(np.percentile(svc, 50)), "p99": float(np.percentile(svc, 99))},
            "service_ms_by_shape": {k: float(np.mean(v)) for k, v in sorted(by_size.items())},
            "errors": sorted({r.error for r in reqs if r.error})[:5]}


def run_stream(rate, n, precision, budget_gb, workers=2, sizes=SIZES, seed=0, log=None, policy="sjf"):
    """Poisson stream of n requests (mean `rate` per second); sizes uniform over
    `sizes`, boundary conditions half open, half periodic."""
    rng = random.Random(seed)
    srv = Server(int(budget_gb * 1e9), workers, policy=policy)
    reqs = []
    t0 = time.perf_counter()
    for i in range(n):
        L = rng.choice(sizes)
        bc = W.OPEN if rng.random() < 0.5 else W.PERIODIC
        r = standard_request(i, L, bc, precision, seed=1000 + i, eps=rng.uniform(0.2, 0.4))
        srv.submit(r)
        reqs.append(r)
        if log and i % 10 == 0:
            log(f"[serve] submitted {i + 1}/{n}")
        time.sleep(rng.expovariate(rate))
    while True:                         # progress while the queue drains
        with srv.lock:
            nd = len(srv.done)
        if nd >= n:
            break
        if log:
            log(f"[serve] completed {nd}/{n}")
        time.sleep(5.0)
    srv.drain()
    out = summarize(reqs, t0)
    out.update({"offered_rate_per_s": rate, "precision": "mixed" if precision == W.MIXED else "fp64",
                "budget_gb": budget_gb, "workers": workers, "sizes": list(sizes), "policy": policy,
                "evicted_contexts": srv.adm.evicted})
    srv.stop()
    return out, reqs


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="mixed")
    ap.add_argument("--budget-gb", type=float, default=260.0)
    ap.add_argument("--workers", type=int, default=2)
    ap.add_argument("--sizes", default=",".join(map(str, SIZES)))
    ap.add_argument("--policy", choices=("sjf", "fifo"), default="sjf",
                    help="sjf: least predicted cost first, with ageing; fifo: arrival order")
    args = ap.parse_args()
    sizes = tuple(int(s) for s in args.sizes.split(","))
    prec = W.MIXED if args.precision == "mixed" else W.FP64
    out, _ = run_stream(args.rate, args.requests, prec, args.budget_gb, args.workers, sizes,
                        log=lambda *a: print(*a, file=sys.stderr, flush=True), policy=args.policy)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
