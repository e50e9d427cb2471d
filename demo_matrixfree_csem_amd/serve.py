"""Serving loop for survey solves on one GPU.

Requests (a model, frequencies, sources, receivers) arrive over time.  The
engine keeps every model it has built resident on the device (FP64 and FP32
coefficients, airwave multipliers, and a pool of solve slots holding the
Krylov vectors), queues each request's right-hand sides (source x
frequency pairs) and runs up to `workers` solves at once, each on its own
stream, each taking up to `rhs_per_solve` waiting right-hand sides of one
model, from one request or several: continuous batching -- a solve starts
as soon as a worker and enough device memory are free while the others keep
running, and solves of different requests and models overlap on the device.
Models are built on a separate thread, so a new model's setup does not
hold up solves on resident ones; the host reads the input files and makes
the 1-D tables, and the device builds the cells, the edge-averaged
conductivities and the airwave multipliers (csem_model_build_device).

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
tol = 1e-8/f as the reference's driver sets it (main.c:104-118).  A
right-hand side's arithmetic is the same in any batch (the FP32 kernels
give the same values alone and batched, and every scalar and convergence
test is per right-hand side) and does not depend on what else runs on the
device, so a request gives the same receivers alone and in the mix.

    python -m demo_matrixfree_csem_amd.serve [--requests N] [--rate R] [--full]
    python -m demo_matrixfree_csem_amd.serve --requests-file reqs.jsonl --out DIR

prints one JSON line: p50/p99 latency, solves/s, the budget and the peaks.
With --requests-file the requests come from a file (REQUEST_SCHEMA below)
and each request's receivers are written under DIR/<id>/ in the
reference's format (emf_0001.txt, one file per source, float32 as
write_data.c:24-34 prints them) and in double precision
(emf_0001_f64.txt, %.17e), with a result.json of latency, iterations and
true residuals.
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
    """One survey model on the device: the input files read and the 1-D
    tables made on the host (under _HOST), the cells, conductivities and
    airwave multipliers built on the device (csem_model_build_device)."""

    def __init__(self, survey, workdir):
        surveys.write_files(workdir, survey)
        with _HOST:
            self.ses = Session(survey.argv(workdir), device=False, setup="device")
        self.model = self.ses.built
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
        """One batch solve: `items` are (request, source index, frequency)
        right-hand sides of one model, from one request or several.
        Returns per item the receivers in double precision (what
        extract_emf rounds to float), the iterations and the true relative
        residual."""
        rm, items, requests = job
        bs = []
        with _HOST:
            for i, s, f in items:
                rm.ses.set_freq(f)  # the finite-volume source scales with i w mu0 (inject_extract.c:51)
                bs.append(rm.ses.inject_at(*requests[i][3][s]))
        fs = [f for i, s, f in items]
        xs, its, rel = self._solve(self.io[wid], rm, bs, fs, [solve_tol(f) for f in fs])
        with _HOST:
            rec = []
            for (i, s, f), x in zip(items, xs):
                rm.ses.set_freq(f)
                rec.append(rm.ses.extract_at(x, requests[i][4], double=True))
        return rec, its, rel

    def _estimate(self, survey):
        """Device bytes a model of `survey` will take: its FP64 arrays (about
        160 bytes per unknown of the full grid -- conductivities, LU-SGS
        records and Smith cache in the sweeps' diagonal-major layout, which
        gives every diagonal a whole wave's slots, about twice the cells,
        pre-pass buffer -- and 80 per airwave transform point: two
        multipliers, three work planes) plus half the FP64 arrays for the
        FP32 copy, times the largest ratio of taken to estimated bytes seen
        so far.  Returns (estimate, unscaled estimate)."""
        nx, ny, nz = (len(a) - 1 for a in (survey.model.x1node, survey.model.x2node, survey.model.x3node))
        nb = int(survey.flags.get("nb", 10))
        n1, n2, n3f = nx + 2 * nb, ny + 2 * nb, nz + 2 * nb
        fE = n1 * (n2 + 1) * (n3f + 1) + (n1 + 1) * n2 * (n3f + 1) + (n1 + 1) * (n2 + 1) * n3f
        nfft = int(survey.flags.get("n1fft", 1024)) if int(survey.flags.get("airbc", 0)) else 0
        base = 1.5 * (160 * fE + 80 * nfft * nfft)
        return int(self.est_scale * base), base

    def run(self, requests, arrivals):
        """Serve `requests` (serve_request tuples) arriving `arrivals`
        seconds after the start.  Returns (results, stats); results[i] maps
        (source index, frequency) -> (receivers in double precision,
        complex128 (channels, nrec): the values extract_emf rounds to float,
        inner iterations, true relative residual) and holds the latency.
        A solve takes up to rhs_per_solve right-hand sides of one model,
        the first waiting one and the next ones on that model from any
        request.

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
        nxt, nsolves, nbatches, nshared = 0, 0, 0, 0
        setup_s = []           # seconds of each model's setup (input files, host tables, device build)
        peak_acc, peak_dev = 0, 0
        self.reserved = 0
        pool = cf.ThreadPoolExecutor(self.workers)
        builder = cf.ThreadPoolExecutor(1)

        def setup(survey):
            self.nbuilt += 1
            t = time.time()
            rm = ResidentModel(survey, os.path.join(self.work, "m%d" % self.nbuilt))
            setup_s.append(time.time() - t)
            return rm

        def enqueue(i, rm):
            key, build, freqs, srcs, _ = requests[i]
            rm.busy += 1
            pairs = [(s, f) for s in range(len(srcs)) for f in freqs]
            pending[i] = len(pairs)
            waiting.extend((i, rm, s, f) for s, f in pairs)

        def next_batch():
            """The first waiting right-hand side and, up to rhs_per_solve,
            the next ones on the same model in arrival order, from any
            request (cross-request batching)."""
            rm = waiting[0][1]
            take = [w for w in range(len(waiting)) if waiting[w][1] is rm][:self.R]
            return rm, take

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
                    rm, take = next_batch()
                    nb = len(take)
                    need = rm.slot_bytes(nb) + max(0, 2 * nb * rm.n * 16 - 2 * self.io[free_workers[0]].cap)
                    if not self._evict_for(need, running):
                        break
                    items = [(waiting[w][0], waiting[w][2], waiting[w][3]) for w in take]
                    for w in reversed(take):
                        waiting.pop(w)
                    w = free_workers.pop(0)
                    running[w] = need
                    peak_acc = max(peak_acc, self._used(running))
                    nbatches += 1
                    nshared += len({it[0] for it in items}) > 1
                    fu = pool.submit(self._work, w, (rm, items, requests))
                    fu.add_done_callback(lambda fu, w=w, rm=rm, items=items: done.put((w, rm, items, fu)))
                inflight = any(b["stage"] != "ready" for b in building.values())
                blocked = waiting or any(b["stage"] == "ready" for b in building.values())
                if blocked and not running and not inflight and all(m.busy for m in self.models.values()):
                    # nothing runs, nothing builds and nothing can be evicted
                    raise RuntimeError("a model or a solve does not fit the budget of %d bytes" % self.budget)
                try:
                    w, rm, items, fu = done.get(timeout=0.002)
                except queue.Empty:
                    continue
                peak_dev = max(peak_dev, _device_used() - base_dev)
                rec, its, rel = fu.result()  # re-raises a worker's error
                del running[w]
                free_workers.append(w)
                nsolves += len(items)
                for (i, s, f), r, it, rr in zip(items, rec, its, rel):
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
                 "batches": nbatches, "cross_request_batches": nshared, "fp64_continued": self.fp64_continued,
                 "model_setup_s": [round(v, 3) for v in setup_s]}
        return results, stats

    def close(self):
        for m in self.models.values():
            m.close()
        self.models = {}
        for b in self.io:
            b.free()
        shutil.rmtree(self.work, ignore_errors=True)


# ------------------------------------------------------------ request files
REQUEST_SCHEMA = """One request per line of a JSON-lines file (or a JSON list):

  {"id": "r0001",                      # names the output directory
   "arrival_s": 0.0,                   # arrival time after the start
   "model": {"files": {"fx1": "...", "fx2": "...", "fx3": "...",
                       "frho11": "...", "frho22": "...", "frho33": "..."},
             "flags": {"nb": 10, "airbc": 1, "fvm": 0, "lextend": 40000,
                       "n1fft": 1024, "chrec": "Ex,Ey,Ez"}}
   "freqs": [0.25, 0.75],              # Hz
   "sources": [[x, y, z], ...],        # Ex dipoles
   "receivers": [[x, y, z], ...]}

"files" are the reference's model inputs (emf_init_free.c:77-150: float32
node coordinates, n + 1 per axis; float32 cell resistivities, nx ny nz, x
fastest), relative paths taken from the request file's directory.  Instead
of "files" a model may be {"synthetic": {"grid": [nx, ny, nz], "kind":
"layered" | "grf", "seed": 1001, "median": 2.0}} (surveys.marine).  Flags
default to those of the reference's examples (nb 10, airbc 1, fvm 0, 1024^2
transform, channels Ex,Ey,Ez); requests with the same model inputs and
flags share one resident model."""

_MODEL_FLAGS = {"nb": 10, "airbc": 1, "fvm": 0, "lextend": None, "n1fft": 1024, "chrec": "Ex,Ey,Ez"}


def read_model_files(paths):
    """A surveys.Model from the reference's binary model inputs."""
    x = [np.fromfile(paths[k], np.float32) for k in ("fx1", "fx2", "fx3")]
    nx, ny, nz = (len(a) - 1 for a in x)
    if min(nx, ny, nz) < 1:
        raise ValueError("node files need at least two coordinates each")
    rho = []
    for k in ("frho11", "frho22", "frho33"):
        r = np.fromfile(paths[k], np.float32)
        if r.size != nx * ny * nz:
            raise ValueError("%s holds %d values, the grid has %d cells" % (paths[k], r.size, nx * ny * nz))
        rho.append(r.reshape(nz, ny, nx))
    return surveys.Model(x[0], x[1], x[2], *rho)


def request_from_dict(d, base_dir="."):
    """(id, arrival, serve_request tuple) of one request-file entry."""
    rid = str(d["id"])
    freqs = sorted(float(np.float32(f)) for f in d["freqs"])
    srcs = [tuple(float(v) for v in p[:3]) for p in d["sources"]]
    recs = [tuple(float(v) for v in p[:3]) for p in d["receivers"]]
    if not freqs or not srcs or not recs:
        raise ValueError("request %s: freqs, sources and receivers must not be empty" % rid)
    spec = d["model"]
    fl = dict(_MODEL_FLAGS)
    fl.update(spec.get("flags", {}))
    flags = dict(fvm=int(fl["fvm"]), airbc=int(fl["airbc"]), preco=1, algopt=1, tol=1e-8, niter=2000,
                 nb=int(fl["nb"]), chsrc="Ex", chrec=str(fl["chrec"]), n1fft=int(fl["n1fft"]),
                 n2fft=int(fl["n1fft"]))
    if fl["lextend"] is not None:  # padding length (default 40 km, emf_init)
        flags["lextend"] = float(fl["lextend"])
    fkey = tuple(sorted(flags.items()))
    if "files" in spec:
        paths = {k: os.path.abspath(os.path.join(base_dir, spec["files"][k]))
                 for k in ("fx1", "fx2", "fx3", "frho11", "frho22", "frho33")}
        key = ("files", tuple(paths[k] for k in sorted(paths)), fkey)

        def model():
            return read_model_files(paths)
        name = "files_" + os.path.basename(paths["frho11"])
    else:
        sy = spec["synthetic"]
        g = tuple(int(v) for v in sy["grid"])
        kind, seed, median = str(sy.get("kind", "layered")), int(sy.get("seed", 1000)), float(sy.get("median", 2.0))
        key = ("synthetic", g, kind, seed, median, fkey)

        def model():
            return surveys.marine(seed, g[0], g[1], g[2], kind=kind, median=median)
        name = "synthetic_%dx%dx%d_%s%d" % (g + (kind, seed))

    def build():
        return surveys.Survey(name, model(), surveys.Acquisition(surveys._acq_table(srcs[:1]),
                                                                 surveys._acq_table(recs)), freqs, dict(flags))
    return rid, float(d.get("arrival_s", 0.0)), (key, build, freqs, srcs, recs)


def load_requests(path):
    """[(id, arrival, request)] from a request file (REQUEST_SCHEMA)."""
    with open(path) as f:
        text = f.read()
    t = text.lstrip()
    entries = json.loads(t) if t.startswith("[") else [json.loads(l) for l in text.splitlines() if l.strip()]
    base = os.path.dirname(os.path.abspath(path))
    out = [request_from_dict(d, base) for d in entries]
    ids = [r[0] for r in out]
    if len(set(ids)) != len(ids):
        raise ValueError("request ids must be unique")
    return out


def write_results(outdir, ids, requests, results):
    """Each request's receivers under outdir/<id>/: per source s
    emf_%04d.txt (s + 1) in the reference's format (write_data.c:24-34:
    channel, then frequency, then receiver; float32 printed with %e) and
    emf_%04d_f64.txt with the double values (%.17e), plus result.json."""
    for rid, req, res in zip(ids, requests, results):
        key, build, freqs, srcs, recs = req
        d = os.path.join(outdir, rid)
        os.makedirs(d, exist_ok=True)
        info = {"id": rid, "latency_s": round(res["latency"], 4), "solves": []}
        for s in range(len(srcs)):
            rows32, rows64 = [], []
            nch = res["receivers"][(s, freqs[0])][0].shape[0]
            for ch in range(nch):
                for fi, f in enumerate(freqs):
                    r = res["receivers"][(s, f)][0][ch]
                    for ir in range(len(recs)):
                        v32 = np.complex64(r[ir])
                        rows32.append("%d \t %d \t %d \t %d \t %e \t %e" % (s + 1, ir + 1, ch + 1, fi + 1,
                                                                             v32.real, v32.imag))
                        rows64.append("%d \t %d \t %d \t %d \t %.17e \t %.17e" % (s + 1, ir + 1, ch + 1, fi + 1,
                                                                                   r[ir].real, r[ir].imag))
            head = "iTx \t iRx    ichrec  ifreq \t emf_real \t emf_imag\n"
            for suffix, rows in (("", rows32), ("_f64", rows64)):
                with open(os.path.join(d, "emf_%04d%s.txt" % (s + 1, suffix)), "w") as f:
                    f.write(head + "\n".join(rows) + "\n")
            for f in freqs:
                _, it, rr = res["receivers"][(s, f)]
                info["solves"].append({"source": s + 1, "freq": f, "iterations": it, "true_relres": rr})
        with open(os.path.join(d, "result.json"), "w") as f:
            json.dump(info, f, indent=1)


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
    ap.add_argument("--requests-file", default=None, help="requests from a file (see REQUEST_SCHEMA); their "
                    "arrival_s replace the Poisson arrivals")
    ap.add_argument("--out", default=None, help="with --requests-file: write each request's receivers here")
    args = ap.parse_args()
    ids = None
    if args.requests_file:
        loaded = load_requests(args.requests_file)
        ids = [r[0] for r in loaded]
        arrivals = [r[1] for r in loaded]
        reqs = [r[2] for r in loaded]
    else:
        if args.full:
            reqs = [surveys.serve_request(args.seed * 1000 + i) for i in range(args.requests)]
        else:
            reqs = quick_requests(args.requests, args.seed)
        arrivals = poisson_arrivals(len(reqs), args.rate, args.seed)
    budget = None if args.budget_gib is None else int(args.budget_gib * 2 ** 30)
    eng = Engine(budget_bytes=budget, workers=args.workers, rhs_per_solve=args.rhs)
    try:
        if not args.cold:
            eng.preload(reqs)
        results, stats = eng.run(reqs, arrivals)
    finally:
        eng.close()
    if ids is not None and args.out:
        write_results(args.out, ids, reqs, results)
    stats["rate_per_s"] = None if args.requests_file else args.rate
    stats["warm"] = not args.cold
    print(json.dumps(stats), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
