"""Serving loop for survey solves on one GPU.

Requests (a model, frequencies, sources, receivers) arrive over time.  The
engine keeps every model it has built resident on the device (FP64 and FP32
coefficients, airwave multipliers, and a pool of solve slots holding the
Krylov vectors), splits each request into solves of up to `rhs_per_solve`
right-hand sides (source x frequency pairs) and runs up to `workers` solves
at once, each on its own stream: continuous batching -- a solve starts as
soon as a worker and enough device memory are free while the others keep
running, and solves of different requests and models overlap on the device.
Models are built on a separate thread, so a new model's host setup does not
hold up solves on resident ones.

Admission is by HBM footprint.  The engine charges every resident model
what it holds on the device (csem_model_bytes + csem_mixed_bytes, the
latter including its pooled slots; an FP32 copy not yet made is charged at
half the FP64 model), a model being uploaded its estimate (from its grid and
transform sizes, scaled by the largest ratio of taken to estimated bytes seen
so far), every worker its I/O
buffers, and every running solve one more slot (csem_mixed_slot_bytes; an
upper bound, as a pooled slot is usually reused).  A model is uploaded and a
solve admitted only if the total stays within the budget; models with no
request in flight are evicted, least recently used first, to make room.

Every solve is csem_mixed_solve (FP32 LU-SGS + F-BiCGStab inside FP64
iterative refinement) to the true FP64 residual |b - A x| / |b| < tol,
tol = 1e-8/f as the reference's driver sets it (main.c:104-118).  Each
right-hand side's arithmetic depends only on its batch, not on what else
runs on the device, so a request gives the same receivers alone and in
the mix.

    python -m demo_matrixfree_csem_amd.serve [--requests N] [--rate R] [--full]

prints one JSON line: p50/p99 latency, solves/s, the budget and the peaks.
"""
import argparse
import concurrent.futures as cf
import ctypes as C
import json
import os
import queue
import shutil
import sys
import tempfile
import threading
import time

import numpy as np

from . import _lib, surveys
from .session import Session, freq_consts

# The reference's host entry points (initargs, emf_init, inject_source,
# extract_emf ...) and the airwave tables share process-global state.
_HOST = threading.Lock()
FP64_NITER = 4000  # iteration budget of an FP64 continuation


def solve_tol(freq):
    """tol/f in single precision, as main.c:104-118 forms it (tol = 1e-8)."""
    return float(np.float32(np.float32(1e-8) / np.float32(freq)))


def fp64_continue(model, freq, dx, db, tol, rel, niter, lock):
    """Finish a right-hand side that the mixed solve stopped short of tol
    on (its FP32 inner iterations stagnate on an ill-conditioned model):
    the FP64 device F-BiCGStab, the reference's solver, from the mixed
    solution in dx (device), asked for the remaining reduction tol / rel
    (its tolerance is relative to the residual it starts from,
    solver.c:115, 158-159).  The FP64 solver works in the model's own
    buffers, so calls on one model hold `lock`.  Returns (iterations, the
    solver's estimate of the final relative residual)."""
    L = _lib.lib()
    hist = np.zeros(niter + 2)
    with lock:
        it = L.csem_bicgstab(model, C.byref(freq_consts(freq)), dx, db, niter, min(0.5, tol / rel), 1,
                             hist.ctypes.data, None)
    _lib.check(it, "csem_bicgstab")
    return it, rel * float(hist[it + 1] if it < niter else hist[niter - 1])


class ResidentModel:
    """One survey model on the device (host setup under _HOST)."""

    def __init__(self, survey, workdir):
        L = _lib.lib()
        surveys.write_files(workdir, survey)
        with _HOST:
            self.ses = Session(survey.argv(workdir), device=False)
            m = C.c_void_p()
            _lib.check(L.csem_model_from_emf(C.byref(self.ses.emf), C.byref(m)), "csem_model_from_emf")
        self.model = m
        self.n = self.ses.n
        self.busy = 0          # requests in flight on this model
        self.last_used = time.time()
        self.fp64_lock = threading.Lock()  # the FP64 solver's buffers are the model's

    def bytes(self):
        L = _lib.lib()
        fp64 = int(L.csem_model_bytes(self.model))
        mixed = int(L.csem_mixed_bytes(self.model))
        return fp64 + (mixed if mixed > 0 else fp64 // 2)

    def slot_bytes(self, R):
        return int(_lib.lib().csem_mixed_slot_bytes(self.model, R))

    def close(self):
        if self.model:
            _lib.lib().csem_model_destroy(self.model)
            self.model = None


class _IO:
    """A worker's device buffers for b and x, grown when a larger solve
    comes (freeing device memory synchronises the device, so they are not
    freed per solve)."""

    def __init__(self):
        self.db, self.dx, self.cap = C.c_void_p(), C.c_void_p(), 0

    def ensure(self, nbytes):
        L = _lib.lib()
        if nbytes <= self.cap:
            return
        self.free()
        _lib.check(L.csem_malloc(C.byref(self.db), nbytes), "csem_malloc")
        _lib.check(L.csem_malloc(C.byref(self.dx), nbytes), "csem_malloc")
        self.cap = nbytes

    def free(self):
        L = _lib.lib()
        if self.cap:
            L.csem_free(self.db)
            L.csem_free(self.dx)
        self.db, self.dx, self.cap = C.c_void_p(), C.c_void_p(), 0


def _device_used():
    free, total = C.c_long(), C.c_long()
    _lib.check(_lib.lib().csem_mem_info(C.byref(free), C.byref(total)), "csem_mem_info")
    return total.value - free.value


class Engine:
    """Continuous-batching solve server over resident models."""

    def __init__(self, budget_bytes=None, workers=3, rhs_per_solve=1, opts=None):
        L = _lib.lib()
        free, total = C.c_long(), C.c_long()
        _lib.check(L.csem_mem_info(C.byref(free), C.byref(total)), "csem_mem_info")
        self.budget = int(0.85 * free.value) if budget_bytes is None else int(budget_bytes)
        self.workers = workers
        self.R = rhs_per_solve
        self.opts = opts or _lib.MixedOpts.default()
        self.models = {}
        self.io = [_IO() for _ in range(workers)]
        self.work = tempfile.mkdtemp(prefix="csem_serve_")
        self.nbuilt = 0
        self.reserved = 0          # estimated bytes of models being uploaded
        self.est_scale = 1.0       # largest ratio of a model's bytes to its estimate
        self.evictions = 0
        self.fp64_continued = 0    # right-hand sides the FP64 solver finished
        self._count_mu = threading.Lock()

    # ------------------------------------------------------ accounting
    def _used(self, running):
        return (sum(m.bytes() for m in self.models.values()) + sum(b.cap * 2 for b in self.io) +
                sum(need for need in running.values()) + self.reserved)

    def _evict_for(self, need, running):
        """Evict idle models, least recently used first, until `need` more
        bytes fit the budget.  Returns whether they fit."""
        while self._used(running) + need > self.budget:
            idle = [(m.last_used, k) for k, m in self.models.items() if m.busy == 0]
            if not idle:
                return False
            k = min(idle)[1]
            self.models.pop(k).close()
            self.evictions += 1
        return True

    def preload(self, requests):
        """Build the models of `requests` ahead of time (a warm server)."""
        for key, build, *_ in requests:
            if key not in self.models:
                self.nbuilt += 1
                survey = build()
                rm = ResidentModel(survey, os.path.join(self.work, "m%d" % self.nbuilt))
                self.est_scale = max(self.est_scale, rm.bytes() / self._estimate(survey)[1])
                self.models[key] = rm
        if self._used({}) > self.budget:
            raise RuntimeError("the preloaded models exceed the budget of %d bytes" % self.budget)

    # ---------------------------------------------------------- solves
    def _solve(self, io, rm, bs, freqs, tols):
        L = _lib.lib()
        R, n = len(bs), rm.n
        io.ensure(R * n * 16)
        fr = (_lib.FreqT * R)(*[freq_consts(f) for f in freqs])
        hb = np.ascontiguousarray(np.concatenate(bs), np.complex128)
        _lib.check(L.csem_memcpy_h2d(io.db, hb.ctypes.data, R * n * 16, None), "upload")
        t = np.ascontiguousarray(tols, np.float64)
        its, rel = np.zeros(R, np.int32), np.zeros(R)
        nc = L.csem_mixed_solve(rm.model, R, fr, io.dx, io.db, t.ctypes.data, C.byref(self.opts),
                                its.ctypes.data, rel.ctypes.data, None)
        _lib.check(nc, "csem_mixed_solve")
        for r in range(R):
            if not rel[r] < tols[r]:
                off = r * n * 16
                it, rel[r] = fp64_continue(rm.model, freqs[r], C.c_void_p(io.dx.value + off),
                                           C.c_void_p(io.db.value + off), tols[r], rel[r], FP64_NITER, rm.fp64_lock)
                its[r] += it
                with self._count_mu:
                    self.fp64_continued += 1
        hx = np.zeros(R * n, np.complex128)
        _lib.check(L.csem_memcpy_d2h(hx.ctypes.data, io.dx, R * n * 16, None), "download")
        return [hx[r * n:(r + 1) * n] for r in range(R)], its, rel

    def _work(self, wid, job):
        i, rm, pairs, req = job
        key, build, freqs, srcs, recs = req
        bs = []
        with _HOST:
            for s, f in pairs:
                rm.ses.set_freq(f)  # the finite-volume source scales with i w mu0 (inject_extract.c:51)
                bs.append(rm.ses.inject_at(*srcs[s]))
        fs = [f for s, f in pairs]
        xs, its, rel = self._solve(self.io[wid], rm, bs, fs, [solve_tol(f) for f in fs])
        with _HOST:
            rec = []
            for (s, f), x in zip(pairs, xs):
                rm.ses.set_freq(f)
                rec.append(rm.ses.extract_at(x, recs))
        return rec, its, rel

    def _estimate(self, survey):
        """Device bytes a model of `survey` will take: its FP64 arrays (about
        96 bytes per unknown of the full grid -- conductivities, LU-SGS
        records, Smith cache, pre-pass buffer -- and 80 per airwave
        transform point: two multipliers, three work planes) plus half that for the
        FP32 copy, times the largest ratio of taken to estimated bytes seen
        so far.  Returns (estimate, unscaled estimate)."""
        nx, ny, nz = (len(a) - 1 for a in (survey.model.x1node, survey.model.x2node, survey.model.x3node))
        nb = int(survey.flags.get("nb", 10))
        n1, n2, n3f = nx + 2 * nb, ny + 2 * nb, nz + 2 * nb
        fE = n1 * (n2 + 1) * (n3f + 1) + (n1 + 1) * n2 * (n3f + 1) + (n1 + 1) * (n2 + 1) * n3f
        nfft = int(survey.flags.get("n1fft", 1024)) if int(survey.flags.get("airbc", 0)) else 0
        base = 1.5 * (96 * fE + 80 * nfft * nfft)
        return int(self.est_scale * base), base

    def run(self, requests, arrivals):
        """Serve `requests` (serve_request tuples) arriving `arrivals`
        seconds after the start.  Returns (results, stats); results[i] maps
        (source index, frequency) -> (receivers complex64 (channels, nrec),
        inner iterations, true relative residual) and holds the latency.

        A request's model is made resident first (its survey built and its
        host setup and upload run on the builder thread, the upload only
        once the model's estimated bytes fit the budget); its solves then
        queue for the workers, first come first served."""
        t0 = time.time()
        base_dev = _device_used()
        results = [None] * len(requests)
        done = queue.Queue()
        free_workers = list(range(self.workers))
        running = {}           # worker -> charged bytes
        pending = {}           # request -> solves left
        recs = {i: {} for i in range(len(requests))}
        arrived = []           # requests whose model is not resident yet
        waiting = []           # (request, model, pairs) ready to admit
        building = {}          # model key -> {"fut", "survey", "est", "units", "stage"}
        order = sorted(range(len(requests)), key=lambda i: arrivals[i])
        nxt, nsolves = 0, 0
        peak_acc, peak_dev = 0, 0
        self.reserved = 0
        pool = cf.ThreadPoolExecutor(self.workers)
        builder = cf.ThreadPoolExecutor(1)

        def setup(survey):
            self.nbuilt += 1
            return ResidentModel(survey, os.path.join(self.work, "m%d" % self.nbuilt))

        def enqueue(i, rm):
            key, build, freqs, srcs, _ = requests[i]
            rm.busy += 1
            pairs = [(s, f) for s in range(len(srcs)) for f in freqs]
            chunks = [pairs[q:q + self.R] for q in range(0, len(pairs), self.R)]
            pending[i] = len(chunks)
            waiting.extend((i, rm, c) for c in chunks)

        try:
            while nxt < len(order) or arrived or waiting or running:
                now = time.time() - t0
                while nxt < len(order) and arrivals[order[nxt]] <= now:
                    i = order[nxt]
                    nxt += 1
                    arrived.append(i)
                    key = requests[i][0]
                    if key not in self.models and key not in building:
                        building[key] = {"fut": builder.submit(requests[i][1]), "stage": "survey"}
                # builds that finished a stage
                for key, b in list(building.items()):
                    if b["fut"] is None or not b["fut"].done():
                        continue
                    if b["stage"] == "survey":
                        b["survey"] = b["fut"].result()
                        b["est"], b["units"] = self._estimate(b["survey"])
                        b["stage"], b["fut"] = "ready", None
                    elif b["stage"] == "setup":
                        rm = b["fut"].result()
                        self.reserved -= b["est"]
                        self.est_scale = max(self.est_scale, rm.bytes() / b["units"])
                        self.models[key] = rm
                        del building[key]
                # requests whose model is resident queue their solves
                still = []
                for i in arrived:
                    rm = self.models.get(requests[i][0])
                    if rm is not None:
                        rm.last_used = time.time()
                        enqueue(i, rm)
                    else:
                        still.append(i)
                arrived = still
                # upload the next ready model, in arrival order, when it fits
                for i in arrived:
                    b = building.get(requests[i][0])
                    if b is None or b["stage"] != "ready":
                        continue
                    if not self._evict_for(b["est"], running):
                        break
                    self.reserved += b["est"]
                    b["stage"], b["fut"] = "setup", builder.submit(setup, b.pop("survey"))
                    break
                # admit solves first come first served while a worker and the memory are free
                while waiting and free_workers:
                    i, rm, c = waiting[0]
                    need = rm.slot_bytes(len(c)) + max(0, 2 * len(c) * rm.n * 16 - 2 * self.io[free_workers[0]].cap)
                    if not self._evict_for(need, running):
                        break
                    waiting.pop(0)
                    w = free_workers.pop(0)
                    running[w] = need
                    peak_acc = max(peak_acc, self._used(running))
                    fu = pool.submit(self._work, w, (i, rm, c, requests[i]))
                    fu.add_done_callback(lambda fu, w=w, i=i, rm=rm, c=c: done.put((w, i, rm, c, fu)))
                inflight = any(b["stage"] != "ready" for b in building.values())
                blocked = waiting or any(b["stage"] == "ready" for b in building.values())
                if blocked and not running and not inflight and all(m.busy for m in self.models.values()):
                    # nothing runs, nothing builds and nothing can be evicted
                    raise RuntimeError("a model or a solve does not fit the budget of %d bytes" % self.budget)
                try:
                    w, i, rm, c, fu = done.get(timeout=0.002)
                except queue.Empty:
                    continue
                peak_dev = max(peak_dev, _device_used() - base_dev)
                rec, its, rel = fu.result()  # re-raises a worker's error
                del running[w]
                free_workers.append(w)
                nsolves += len(c)
                for (s, f), r, it, rr in zip(c, rec, its, rel):
                    recs[i][(s, f)] = (r, int(it), float(rr))
                pending[i] -= 1
                if pending[i] == 0:
                    rm.busy -= 1
                    rm.last_used = time.time()
                    results[i] = {"receivers": recs[i], "latency": time.time() - t0 - arrivals[i]}
        finally:
            pool.shutdown(wait=True)
            builder.shutdown(wait=True)
        wall = time.time() - t0
        lat = np.array([r["latency"] for r in results])
        stats = {"requests": len(requests), "solves": nsolves, "wall_s": round(wall, 3),
                 "solves_per_s": round(nsolves / wall, 4),
                 "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
                 "p99_latency_s": round(float(np.percentile(lat, 99)), 3),
                 "budget_bytes": self.budget, "peak_admitted_bytes": int(peak_acc),
                 "peak_device_bytes": int(peak_dev), "resident_models": len(self.models),
                 "models_built": self.nbuilt, "evictions": self.evictions, "workers": self.workers, "rhs_per_solve": self.R,
                 "fp64_continued": self.fp64_continued}
        return results, stats

    def close(self):
        for m in self.models.values():
            m.close()
        self.models = {}
        for b in self.io:
            b.free()
        shutil.rmtree(self.work, ignore_errors=True)


def poisson_arrivals(n, rate, seed):
    """Arrival times of a Poisson process of `rate` requests per second."""
    rng = np.random.default_rng(seed)
    return [float(t) for t in np.cumsum(rng.exponential(1.0 / rate, n))]


def quick_requests(n, seed):
    """The default stream: 64^3 and 100^3 grids, two models each, 1-3
    frequencies, 1-2 sources."""
    return [surveys.serve_request(seed * 1000 + i, grids=surveys.SERVE_GRIDS[:2], max_freqs=3, max_sources=2,
                                  n_models=2) for i in range(n)]


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--requests", type=int, default=8)
    ap.add_argument("--rate", type=float, default=0.5, help="arrivals per second")
    ap.add_argument("--workers", type=int, default=3)
    ap.add_argument("--rhs", type=int, default=1, help="right-hand sides per solve (1..4)")
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--budget-gib", type=float, default=None, help="HBM budget (default 85%% of free)")
    ap.add_argument("--cold", action="store_true", help="build models on first use instead of ahead of time")
    ap.add_argument("--full", action="store_true",
                    help="the full request distribution (grids up to 180x120x120, 1-6 frequencies, "
                         "1-4 sources)")
    args = ap.parse_args()
    if args.full:
        reqs = [surveys.serve_request(args.seed * 1000 + i) for i in range(args.requests)]
    else:
        reqs = quick_requests(args.requests, args.seed)
    budget = None if args.budget_gib is None else int(args.budget_gib * 2 ** 30)
    eng = Engine(budget_bytes=budget, workers=args.workers, rhs_per_solve=args.rhs)
    try:
        if not args.cold:
            eng.preload(reqs)
        _, stats = eng.run(reqs, poisson_arrivals(len(reqs), args.rate, args.seed))
    finally:
        eng.close()
    stats["rate_per_s"] = args.rate
    stats["warm"] = not args.cold
    print(json.dumps(stats), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
