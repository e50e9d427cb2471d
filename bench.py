#!/usr/bin/env python3
"""Converged solves per second of the single-rank frequency-domain CSEM
solve on one MI355X: the batched mixed-precision path beside the plain FP64
drop-in path.

Workload (one step): the 100^3-cell layered marine model (100 m cells, 8
seeded log-uniform layers under 1000 m of 0.3125 ohm-m seawater, nb=10,
airwave boundary with a 1024^2 transform), one Ex source, three
frequencies 0.25 / 0.75 / 1.25 Hz, each solved from x = 0 to tol = 1e-8/f
with LU-SGS + F-BiCGStab:
  * plain: the reference's solver in complex FP64 on the device, one
    frequency after the other, to its recursive residual (solver.c:157);
  * mixed (the headline value): the three right-hand sides in one batch,
    FP32 inner solves inside FP64 iterative refinement, to the true FP64
    residual |b - A x| / |b| < tol.
The iteration budget is 2000 (the reference itself needs 1073 iterations at
0.25 Hz, beyond the 1000 of its small-model example).  Inputs (model,
right-hand sides) are resident in HBM before the timed region.

After the timed steps, a serving run (demo_matrixfree_csem_amd/serve.py)
feeds a warm engine seeded requests on 64^3 and 100^3 models at Poisson
arrival times and reports its solves/s and p50/p99 request latency.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line.  --gpus above 1 is refused: this path runs on one
GPU (independent solves shard across GPUs without communication).
"""
import argparse
import ctypes as C
import json
import os
import sys
import tempfile
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/s @ tol 1e-8: batched mixed vs plain FP64 GPU; apply HBM TB/s; p99 lat"
HBM_PEAK_GBS = 8000.0


def _traffic(key="bytes_per_launch"):
    """HBM bytes per launch by kernel (or, for the 180x120x120 operator
    runs, by run), from the FETCH_SIZE / WRITE_SIZE counter runs summarised
    in profiles/hbm_traffic.json (absent: {})."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f).get(key, {})


def _roofline(kl, name, copy_gbs=None):
    """Roofline of kernel `name` from the roofline pass's timings `kl`:
    algorithmic bytes per launch over the average launch duration, each
    launch of the solve between its own HIP-event pair on its stream."""
    k = kl.get(name) if kl else None
    if not k or not k["launches"]:
        return None
    gbs = k["bytes_per_launch"] / (k["event_us"] * 1e3)
    tr = _traffic().get(name)
    out = {"kernel": name, "bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": int(tr) if tr else None,
           "bytes_per_launch": int(k["bytes_per_launch"]), "avg_us": k["event_us"], "launches": k["launches"]}
    if copy_gbs:
        out["frac_of_copy"] = round(gbs / copy_gbs, 4)
    return out


# kernels also timed in back-to-back repeat loops (csem_kernel_repeat): the
# kernels of an iteration whose output does not feed their own input
LOOP_KERNELS = ("k_mx_apply", "k_mx_hz0", "k_mx_air_rows", "k_mx_air_cols", "k_mx_air_ghost", "k_mx_lu_pre",
                "k_mx_sgs", "k_mx_unskew", "k_apply", "k_hz0", "k_air_rows_fwd", "k_air_cols", "k_air_ghost",
                "k_lu_pre", "k_sgs_ex<fwd>", "k_sgs_ey<fwd>", "k_sgs_ez<fwd>", "k_sgs_ez<bwd>", "k_sgs_ey<bwd>",
                "k_sgs_ex<bwd>", "k_dot_partial")
LOOP_REPS, LOOP_LOOPS = 10, 24  # 24 launches: four LU-SGS applications' six of each kind


def _loop_stats(st, looped):
    """Per kernel: its launches and event-timed duration per launch and its
    algorithmic bytes per launch; for the kernels in `looped`, the repeat
    loops' time per launch and the bandwidth it gives."""
    from demo_matrixfree_csem_amd import _lib
    out = {}
    for k, (ms, cnt, by) in st.items():
        e = {"launches": cnt, "event_us": round(1e3 * ms / cnt, 3), "bytes_per_launch": int(by / cnt),
             "event_gbs": round(by / cnt / (1e6 * ms / cnt), 1) if ms > 0 else None}
        if k in looped:
            lms, lcnt = _lib.kernel_loop_stats(k)
            if lcnt:
                e["loop_us"], e["loop_launches"] = round(1e3 * lms / lcnt, 3), lcnt
                e["loop_gbs"] = round(e["bytes_per_launch"] / (1e3 * e["loop_us"]), 1)
        out[k] = e
    return out


def roofline_pass(L, model, n, fr, db, dxm, dxp, iters=40, loops=False):
    """Every kernel of one solve, timed per launch.  An eager
    single-right-hand-side mixed solve of exactly `iters` inner iterations
    (tol 0 and no stagnation cut: the right-hand side works throughout),
    then `iters` plain FP64 F-BiCGStab iterations, with the kernel timers
    on: each launch between its own HIP-event pair on its stream.  Measured
    without a profiler, the pair spans the kernel and under 1 us more
    (rocprof's durations of the same launches agree to about 1 %).
    With `loops`, each of the first LOOP_LOOPS launches of every kernel in
    LOOP_KERNELS is also followed by LOOP_REPS launches of it with the same
    arguments, back to back between one event pair: the kernel with its
    inputs warm in the caches (10-25 % faster than in the solve), reported
    beside the per-launch figure, never in place of it.
    Returns ({kernel: stats}, dominant mixed kernel by summed time)."""
    from demo_matrixfree_csem_amd import _lib
    L.csem_sync(None)
    L.csem_kernel_timing(1)
    if loops:
        for k in LOOP_KERNELS:
            _lib.check(L.csem_kernel_repeat(_lib.kernel_id(k), LOOP_REPS, LOOP_LOOPS), "csem_kernel_repeat")
    ko = _lib.MixedOpts.default(max_inner=iters, max_outer=1, check_every=iters, use_graph=0, inner_tol=1e-300,
                                stall_iters=-1)
    its, rel, never = np.zeros(1, np.int32), np.zeros(1), np.zeros(1)
    L.csem_mixed_solve(model, 1, C.byref(fr), dxm, db, never.ctypes.data, C.byref(ko), its.ctypes.data,
                       rel.ctypes.data, None)
    L.csem_sync(None)
    mixed = _lib.kernel_stats()
    L.csem_memset(dxp, 0, n * 16, None)
    hist = np.zeros(iters + 2)
    _lib.check(L.csem_bicgstab(model, C.byref(fr), dxp, db, iters, 0.0, 1, hist.ctypes.data, None), "csem_bicgstab")
    L.csem_sync(None)
    out = _loop_stats(_lib.kernel_stats(), LOOP_KERNELS if loops else ())
    L.csem_kernel_timing(0)
    timed = [k for k in mixed if k.startswith("k_mx_")]
    dominant = max(timed, key=lambda k: mixed[k][0]) if timed else None
    return out, dominant


def _copy_rate(L, reps=10):
    """GB/s of a plain device copy of 1 GiB (k_copy16: 16-byte loads and
    stores), bytes read plus written over the host-timed duration of `reps`
    back-to-back copies: the bandwidth a streaming kernel can reach here."""
    from demo_matrixfree_csem_amd import _lib
    a, b = C.c_void_p(), C.c_void_p()
    for d in (a, b):
        _lib.check(L.csem_malloc(C.byref(d), 2 ** 30), "csem_malloc")
    try:
        for _ in range(2):
            _lib.check(L.csem_copy(b, a, 2 ** 30, None), "csem_copy")
        L.csem_sync(None)
        t0 = time.perf_counter()
        for _ in range(reps):
            _lib.check(L.csem_copy(b, a, 2 ** 30, None), "csem_copy")
        L.csem_sync(None)
        dt = time.perf_counter() - t0
    finally:
        L.csem_free(a)
        L.csem_free(b)
    return 2 * 2 ** 30 * reps / dt / 1e9


def _lusgs_roofline(kl, dims, fp32):
    """One LU-SGS application against the bytes the preconditioner needs
    per unknown: forward Ex 40, Ey 56, Ez 72, backward 88 each (b, sigma,
    the values read, the value written) in complex FP64, half that in
    complex FP32 for the one right-hand side of the roofline pass.  The
    application's time is the sum of its launches' event-timed durations,
    each kernel counted as often per application as the pass launched it:
    FP32 six pre-passes, six sweeps and the two copies of the Ex sweeps'
    values to W and y; FP64 six pre-passes and the six sweeps."""
    n1, n2, n3f = dims
    nc = [n1 * (n3f - 1) * (n2 - 1), n2 * (n3f - 1) * (n1 - 1), n3f * (n2 - 1) * (n1 - 1)]
    per = float(sum(c * (f + 88) for c, f in zip(nc, (40, 56, 72))))
    if fp32:
        napp = kl["k_mx_sgs"]["launches"] / 6.0 if "k_mx_sgs" in kl else 0
        parts = [k for k in ("k_mx_lu_pre", "k_mx_sgs", "k_mx_unskew") if k in kl]
        per /= 2.0
        what = "one FP32 LU-SGS application (one right-hand side): 6 pre-passes + 6 sweeps + 2 copies (Ex)"
    else:
        napp = kl["k_sgs_ex<fwd>"]["launches"] if "k_sgs_ex<fwd>" in kl else 0
        parts = ["k_lu_pre"] + ["k_sgs_%s<%s>" % (c, d) for d in ("fwd", "bwd") for c in ("ex", "ey", "ez")]
        what = "one FP64 LU-SGS application: 6 pre-passes + 6 sweeps"
    if not napp or any(k not in kl for k in parts):
        return None
    us = {k: kl[k]["launches"] / napp * kl[k]["event_us"] for k in parts}
    tot = sum(us.values())
    gbs = per / (tot * 1e3)
    return {"what": what, "bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_application": int(per), "avg_us": round(tot, 3),
            "parts_us": {k: round(v, 3) for k, v in us.items()},
            "launches_per_application": {k: round(kl[k]["launches"] / napp, 2) for k in parts}}


def apply_large(reps=20, copy_gbs=None):
    """The operator on the 180 x 120 x 120 anisotropic model (11.1 M
    unknowns: the FP64 apply moves 443 MB, beyond the 256 MB Infinity
    Cache, so every launch streams from HBM): `reps` launches of the FP64
    operator and of the FP32 one at R = 1 and R = 3, each between its own
    HIP-event pair on its stream.  This replaces the process's active model
    (run it last)."""
    from demo_matrixfree_csem_amd import surveys, _lib
    from demo_matrixfree_csem_amd.session import Session, freq_consts
    L = _lib.lib()
    s = surveys.vti_marine()
    work = tempfile.mkdtemp(prefix="csem_bench_vti_")
    surveys.write_files(work, s)
    ses = Session(s.argv(work))
    model, n = ses.model, ses.n
    rng = np.random.default_rng(7)
    hx = (rng.standard_normal(3 * n) + 1j * rng.standard_normal(3 * n)).astype(np.complex128)
    dx, dy = C.c_void_p(), C.c_void_p()
    for d in (dx, dy):
        _lib.check(L.csem_malloc(C.byref(d), 3 * n * 16), "csem_malloc")
    out = {"workload": "vti_marine 180x120x120 + nb=10 (n_op %d), 0.5 Hz" % n, "n_op": n}
    try:
        _lib.check(L.csem_memcpy_h2d(dx, hx.ctypes.data, 3 * n * 16, None), "upload")
        fr = freq_consts(0.5)
        frs = (_lib.FreqT * 3)(fr, fr, fr)
        h32 = hx.astype(np.complex64)
        _lib.check(L.csem_memcpy_h2d(dy, h32.ctypes.data, 3 * n * 8, None), "upload")
        dy32 = C.c_void_p(dy.value + 3 * n * 8)
        runs = [("fp64", lambda: L.csem_apply(model, C.byref(fr), dx, C.c_void_p(dx.value + n * 16), None), "k_apply")]
        for R in (1, 3):
            runs.append(("fp32_r%d" % R, (lambda R=R: L.csem_apply32(model, R, frs, dy, dy32, None)), "k_mx_apply"))
        for name, fn, kern in runs:
            _lib.check(fn(), name)  # warm (creates the FP32 state on first use)
            L.csem_sync(None)
            L.csem_kernel_timing(1)
            for _ in range(reps):
                _lib.check(fn(), name)
            L.csem_sync(None)
            kl = _loop_stats(_lib.kernel_stats(), ())
            L.csem_kernel_timing(0)
            out[name] = _roofline(kl, kern, copy_gbs)
            tr = _traffic("apply_large_bytes_per_launch").get(name)
            if out[name]:
                out[name]["traffic"] = int(tr) if tr else None
        rows, planes = C.c_int(), C.c_int()
        for R in (1, 3):
            if L.csem_mixed_tuning(model, R, C.byref(rows), C.byref(planes)) == 0:
                out["fp32_r%d_tile" % R] = {"rows": rows.value, "planes": planes.value}
    finally:
        L.csem_free(dx)
        L.csem_free(dy)
        ses.close()
    return out


def _rec_gap(a, b):
    """per channel: max |a - b| over receivers / max |b| of the channel"""
    sc = np.abs(b).max(axis=1)
    return [float("%.2e" % (np.abs(a[c] - b[c]).max() / sc[c])) if sc[c] > 0 else 0.0 for c in range(a.shape[0])]


def aniso_leg(freqs, niter):
    """The anisotropic VTI model (180 x 120 x 120 cells, 11.1 M unknowns,
    six channels) at `freqs`: LU-SGS preconditioned F-BiCGStab and
    FGMRES(10), each in plain FP64 on the device (the reference's solvers,
    one frequency after the other) and in mixed precision (FP32 inner
    F-BiCGStab or FGMRES(10) cycles inside FP64 refinement, one solve per
    frequency on three host threads).  Solves/s count the solves whose true
    FP64 relative residual ends below tol; receivers of each mixed solve are
    compared with the plain FP64 solve of the same method."""
    from demo_matrixfree_csem_amd import surveys, _lib
    from demo_matrixfree_csem_amd.session import Session, freq_consts
    L = _lib.lib()
    s = surveys.vti_marine()
    work = tempfile.mkdtemp(prefix="csem_bench_vti_")
    surveys.write_files(work, s)
    ses = Session(s.argv(work))
    model, n = ses.model, ses.n
    R = len(freqs)
    tols = [float(np.float32(np.float32(1e-8) / np.float32(f))) for f in freqs]
    fr = (_lib.FreqT * R)(*[freq_consts(f) for f in freqs])
    hb = []
    for f in freqs:
        ses.set_freq(f)
        hb.append(ses.inject())
    hb = np.ascontiguousarray(np.concatenate(hb), np.complex128)
    db, dx, dw = C.c_void_p(), C.c_void_p(), C.c_void_p()
    for d, nb in ((db, R * n * 16), (dx, R * n * 16), (dw, n * 16)):
        _lib.check(L.csem_malloc(C.byref(d), nb), "csem_malloc")
    out = {"workload": "vti_marine 180x120x120 + nb=10 (n_op %d), lextend 120 km, LU-SGS, f=%s Hz, tol=1e-8/f"
                       % (n, "/".join("%g" % f for f in freqs)), "n_op": n, "niter": niter}
    xs = {}

    def at(p, r):
        return C.c_void_p(p.value + r * n * 16)

    def true_rel(r):
        L.csem_apply(model, C.byref(fr[r]), at(dx, r), dw, None)
        w = np.zeros(n, np.complex128)
        L.csem_memcpy_d2h(w.ctypes.data, dw, n * 16, None)
        L.csem_sync(None)
        br = hb[r * n:(r + 1) * n]
        return float(np.linalg.norm(br - w) / np.linalg.norm(br))

    def fetch(key):
        h = np.zeros(R * n, np.complex128)
        L.csem_memcpy_d2h(h.ctypes.data, dx, R * n * 16, None)
        L.csem_sync(None)
        xs[key] = h

    try:
        _lib.check(L.csem_memcpy_h2d(db, hb.ctypes.data, R * n * 16, None), "upload")
        for name, solve in (
                ("fp64_fbicgstab", lambda r: L.csem_bicgstab(model, C.byref(fr[r]), at(dx, r), at(db, r), niter,
                                                             tols[r], 1, None, None)),
                ("fp64_fgmres10", lambda r: L.csem_fgmres(model, C.byref(fr[r]), at(dx, r), at(db, r), niter // 10,
                                                          tols[r], 10, 1, None, None))):
            L.csem_memset(dx, 0, R * n * 16, None)
            L.csem_sync(None)
            t0 = time.perf_counter()
            its = []
            for r in range(R):
                its.append(_lib.check(solve(r), name))
            L.csem_sync(None)
            dt = time.perf_counter() - t0
            rel = [true_rel(r) for r in range(R)]
            ok = sum(rel[r] < tols[r] for r in range(R))
            out[name] = {"solves_per_s": round(ok / dt, 4), "wall_s": round(dt, 3), "iterations": its,
                         "true_relres": [float("%.3e" % v) for v in rel], "converged": ok}
            fetch(name)
        for name, kry in (("mixed_fbicgstab", _lib.KRYLOV_FBICGSTAB), ("mixed_fgmres10", _lib.KRYLOV_FGMRES)):
            opts = _lib.MixedOpts.default(max_inner=niter, krylov=kry, restart=10)
            L.csem_memset(dx, 0, R * n * 16, None)
            L.csem_sync(None)
            res = [None] * R
            nxt = [0]
            mu = threading.Lock()

            def worker():
                while True:
                    with mu:
                        r = nxt[0]
                        nxt[0] += 1
                    if r >= R:
                        return
                    itr, rr = np.zeros(1, np.int32), np.zeros(1)
                    tl = np.array([tols[r]])
                    nc = L.csem_mixed_solve(model, 1, C.byref(fr[r]), at(dx, r), at(db, r), tl.ctypes.data,
                                            C.byref(opts), itr.ctypes.data, rr.ctypes.data, None)
                    res[r] = (nc, int(itr[0]), float(rr[0]))

            t0 = time.perf_counter()
            th = [threading.Thread(target=worker) for _ in range(min(3, R))]
            for t in th:
                t.start()
            for t in th:
                t.join()
            L.csem_sync(None)
            dt = time.perf_counter() - t0
            for o in res:
                _lib.check(o[0], name)
            ok = sum(o[2] < tols[r] for r, o in enumerate(res))
            out[name] = {"solves_per_s": round(ok / dt, 4), "wall_s": round(dt, 3),
                         "inner_iterations": [o[1] for o in res], "true_relres": [float("%.3e" % o[2]) for o in res],
                         "converged": ok}
            fetch(name)
        for mixed, plain in (("mixed_fbicgstab", "fp64_fbicgstab"), ("mixed_fgmres10", "fp64_fgmres10")):
            gaps = []
            for r, f in enumerate(freqs):
                ses.set_freq(f)  # the H channels scale by 1 / (i w mu0)
                a = ses.extract_double(xs[mixed][r * n:(r + 1) * n])
                b = ses.extract_double(xs[plain][r * n:(r + 1) * n])
                gaps.append(_rec_gap(a, b))
            out[mixed]["receiver_err_vs_%s" % plain] = gaps
        for m, p in (("mixed_fbicgstab", "fp64_fbicgstab"), ("mixed_fgmres10", "fp64_fgmres10")):
            if out[p]["solves_per_s"] > 0:
                out[m]["speedup_vs_%s" % p] = round(out[m]["solves_per_s"] / out[p]["solves_per_s"], 3)
        out["channels"] = s.flags["chrec"]
        out["receiver_err_what"] = "per frequency, per channel: max |mixed - FP64| / max |FP64| over the 121 receivers"
    finally:
        for d in (db, dx, dw):
            L.csem_free(d)
        ses.close()
    return out


def topo_leg(workers=4, rhs_per_solve=1):
    """The batched topography workload: 8 Ex sources x 6 frequencies on the
    125 x 125 x 100 seafloor-topography model, plus 16 solves on 64^3 and
    96 x 96 x 80 grids of the 64^3 parity model's construction (4 sources x
    2 frequencies each), served by the engine with every request at t = 0:
    up to `rhs_per_solve` right-hand sides of one model per solve (hipGraph
    replays, autotuned tiles), `workers` solves concurrently on their own
    streams, the three models resident (built beforehand).  The defaults
    are the fastest of the (workers, batch) pairs measured on this
    workload: single right-hand sides on four streams."""
    from demo_matrixfree_csem_amd import surveys, serve
    topo = surveys.topo_marine()
    tfreqs = list(topo.freqs)
    trec = [tuple(float(v) for v in r[:3]) for r in topo.acq.receivers]
    reqs = [(("topo", 125, 125, 100), surveys.topo_marine, tfreqs, [tuple(float(v) for v in src[:3])], trec)
            for src in topo.acq.sources]
    for g in ((64, 64, 64), (96, 96, 80)):
        v = surveys.layered_64_variant(*g)
        reqs.append((("layered",) + g, (lambda g=g: surveys.layered_64_variant(*g)), list(v.freqs),
                     [tuple(float(x) for x in src[:3]) for src in v.acq.sources],
                     [tuple(float(x) for x in r[:3]) for r in v.acq.receivers]))
    eng = serve.Engine(workers=workers, rhs_per_solve=rhs_per_solve)
    try:
        t0 = time.perf_counter()
        eng.preload(reqs)
        build_s = time.perf_counter() - t0
        results, st = eng.run(reqs, [0.0] * len(reqs))
    finally:
        eng.close()
    its, worst = {}, 0.0
    for i, rq in enumerate(reqs):
        for (s_i, f), (_, it, rel) in results[i]["receivers"].items():
            its.setdefault(rq[0][0] if rq[0][0] == "topo" else "x".join(map(str, rq[0][1:])), []).append(int(it))
            worst = max(worst, float(rel) / serve.solve_tol(f))
    st.update({"solves_expected": 48 + 16, "model_build_s": round(build_s, 3),
               "inner_iterations_mean": {k: round(float(np.mean(v)), 1) for k, v in its.items()},
               "worst_relres_over_tol": float("%.3e" % worst),
               "workload": "125x125x100 topography (8 Ex sources x 6 frequencies 0.125-1.25 Hz) + 64^3 and "
                           "96x96x80 layered variants (4 sources x 0.25/0.75 Hz each); mixed precision, tol 1e-8/f; "
                           "%d concurrent solves of up to %d right-hand sides" % (workers, rhs_per_solve)})
    return st


def cpu_baseline(survey, freqs, iters, gpu_iters):
    """The oracle (serial C restatement of the reference path) on this
    host: `iters` LU-SGS F-BiCGStab iterations of one solve on one thread,
    scaled to solves/s with the iteration counts the plain GPU path needed
    (the reference is single-threaded, so one core is its CPU path); and,
    beside it, the throughput of one such solve per frequency on as many
    threads at once (ctypes releases the interpreter lock in the C code),
    each frequency's rate scaled by its own GPU iteration count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    m = survey.model
    f = survey.flags

    def make(freq):
        orc = O.Oracle(m.x1node, m.x2node, m.x3node, m.rho11, m.rho22, m.rho33, nb=f["nb"], airbc=f["airbc"],
                       fvm=f.get("fvm", 0), nfft=1024)
        orc.set_freq(freq)
        return orc, orc.inject(0, *survey.acq.sources[0][:3])

    orc, b = make(freqs[0])
    t0 = time.perf_counter()
    orc.fbicgstab(b, iters, 0.0)
    dt = time.perf_counter() - t0
    per_iter = dt / iters
    mean_iters = float(np.mean([i + 1 for i in gpu_iters]))
    del orc
    out = {"value": round(1.0 / (per_iter * mean_iters), 6), "unit": "solves/s", "cores": 1, "kind": "port",
           "sample": "%d LU-SGS F-BiCGStab iterations of the %.2f Hz solve on the same model (serial C "
                     "restatement, 1 thread, %.1f s, %.3f s/iteration), scaled by the plain GPU path's mean of "
                     "%.1f iterations per solve" % (iters, freqs[0], dt, per_iter, mean_iters)}
    # throughput: the frequencies at once, one thread each
    models = [make(fq) for fq in freqs]
    times = [0.0] * len(freqs)

    def run(r):
        t = time.perf_counter()
        models[r][0].fbicgstab(models[r][1], iters, 0.0)
        times[r] = time.perf_counter() - t

    th = [threading.Thread(target=run, args=(r,)) for r in range(len(freqs))]
    t0 = time.perf_counter()
    for t in th:
        t.start()
    for t in th:
        t.join()
    wall = time.perf_counter() - t0
    est = max(times[r] / iters * (gpu_iters[r] + 1) for r in range(len(freqs)))
    out["throughput"] = {"value": round(len(freqs) / est, 6), "unit": "solves/s", "cores": len(freqs),
                         "sample": "the %d frequencies' solves at once, one thread each, %d iterations each "
                                   "(%.1f s wall), each scaled by its plain GPU iteration count; the slowest "
                                   "sets the rate" % (len(freqs), iters, wall)}
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="layered_100")
    ap.add_argument("--cpu-iters", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 path (profiling runs)")
    ap.add_argument("--niter", type=int, default=2000)
    ap.add_argument("--serve-requests", type=int, default=12, help="requests of the serving run (0: none)")
    ap.add_argument("--serve-cold", type=int, default=6,
                    help="requests of the cold serving run over the full grid set (0: none)")
    ap.add_argument("--serve-rate", type=float, default=2.0, help="Poisson arrivals per second of the serving run")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the mixed inner iterations eagerly instead of replaying a captured hipGraph")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the roofline pass (an eager single-right-hand-side mixed solve of 40 inner "
                         "iterations and 40 plain FP64 iterations at the lowest frequency, every launch timed by "
                         "its own event pair) and a 1 GiB device copy of known size (for scaling FETCH_SIZE / "
                         "WRITE_SIZE); prints the pass's JSON instead of the bench line.  The command the "
                         "rocprofv3 summaries and counter runs are taken on")
    ap.add_argument("--apply-only", action="store_true",
                    help="counter runs: only the operator launches on the 180x120x120 model and the copy, then exit")
    ap.add_argument("--no-large", action="store_true", help="skip the 180x120x120 operator measurement")
    ap.add_argument("--aniso-freqs", default="0.75,1.25",
                    help="frequencies of the anisotropic-model leg (both Krylov methods in both precisions), "
                         "comma-separated, 'all' for the workload's six, 'none' to skip")
    ap.add_argument("--aniso-niter", type=int, default=6000)
    ap.add_argument("--no-topo", action="store_true",
                    help="skip the batched topography leg (48 solves + 16 on two smaller grids)")
    ap.add_argument("--legs-only", action="store_true",
                    help="run only the anisotropic and topography legs and print their JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only (this path has no multi-GPU form; independent solves would "
              "shard across GPUs without communication) -- got --gpus %d" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from demo_matrixfree_csem_amd import surveys, _lib
    from demo_matrixfree_csem_amd.session import Session, freq_consts
    L = _lib.lib()

    survey = surveys.GENERATORS[args.workload]()
    work = tempfile.mkdtemp(prefix="csem_bench_")
    surveys.write_files(work, survey)
    ses = Session(survey.argv(work))
    model, n = ses.model, ses.n
    freqs = sorted(ses.freqs)
    R = len(freqs)
    tol0 = np.float32(survey.flags.get("tol", 1e-8))
    tols = [float(np.float32(tol0 / np.float32(f))) for f in freqs]

    # resident inputs: the right-hand sides back to back, and the solutions
    hb = []
    for f in freqs:
        ses.set_freq(f)
        hb.append(ses.inject())
    hb = np.ascontiguousarray(np.concatenate(hb), np.complex128)
    fr = (_lib.FreqT * R)(*[freq_consts(f) for f in freqs])
    db, dxp, dxm, dxs = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    for d in (db, dxp, dxm, dxs):
        _lib.check(L.csem_malloc(C.byref(d), R * n * 16), "csem_malloc")
    _lib.check(L.csem_memcpy_h2d(db, hb.ctypes.data, R * n * 16, None), "upload")
    hist = np.zeros(args.niter + 2)
    tol_arr = np.array(tols)
    its = np.zeros(R, np.int32)
    rel = np.zeros(R)
    opts = _lib.MixedOpts.default(max_inner=args.niter, use_graph=0 if args.no_graph else 1)

    def copy_1gib():
        ca, cb = C.c_void_p(), C.c_void_p()
        for d in (ca, cb):
            _lib.check(L.csem_malloc(C.byref(d), 2 ** 30), "csem_malloc")
        for _ in range(3):
            _lib.check(L.csem_copy(cb, ca, 2 ** 30, None), "csem_copy")
        L.csem_sync(None)
        L.csem_free(ca)
        L.csem_free(cb)

    if args.roofline_only:
        kl, dom = roofline_pass(L, model, n, fr[0], db, dxm, dxp)
        copy_1gib()
        print(json.dumps({"roofline_only": args.workload, "n_op": n, "freq": freqs[0], "dominant": dom,
                          "roofline": _roofline(kl, dom), "copy_bytes_each_way": 2 ** 30, "kernels": kl}),
              flush=True)
        return
    if args.apply_only:
        res = apply_large()
        copy_1gib()
        print(json.dumps({"apply_only": res, "copy_bytes_each_way": 2 ** 30}), flush=True)
        return

    def legs():
        o = {}
        if args.aniso_freqs != "none":
            af = (surveys.vti_marine(2, 8, 8, 8, 3).freqs if args.aniso_freqs == "all"
                  else [float(v) for v in args.aniso_freqs.split(",")])
            print("leg: anisotropic model, f = %s" % af, file=sys.stderr, flush=True)
            o["aniso"] = aniso_leg(af, args.aniso_niter)
        if not args.no_topo:
            print("leg: batched topography workload", file=sys.stderr, flush=True)
            o["topo_batch"] = topo_leg()
        return o

    if args.legs_only:
        ses.close()
        print(json.dumps(legs()), flush=True)
        return

    def plain_step():
        out = []
        for r in range(R):
            xr = C.c_void_p(dxp.value + r * n * 16)
            br = C.c_void_p(db.value + r * n * 16)
            L.csem_memset(xr, 0, n * 16, None)
            it = L.csem_bicgstab(model, C.byref(fr[r]), xr, br, args.niter, tols[r], 1, hist.ctypes.data, None)
            _lib.check(it, "csem_bicgstab")
            out.append(it)
        return out

    # a right-hand side the mixed solve stops short of tol on is finished by
    # the FP64 solver from the mixed solution, inside the timed step, as the
    # serving engine does: every solve counted in the rate reaches tol
    from demo_matrixfree_csem_amd import serve
    fp64_lock = threading.Lock()

    def finish(r, dx, relr):
        if relr < tols[r]:
            return 0, relr
        it, relr = serve.fp64_continue(model, freqs[r], dx, C.c_void_p(db.value + r * n * 16), tols[r], relr,
                                       args.niter, fp64_lock)
        return 1, relr

    def mixed_step():
        nc = L.csem_mixed_solve(model, R, fr, dxm, db, tol_arr.ctypes.data, C.byref(opts), its.ctypes.data,
                                rel.ctypes.data, None)
        _lib.check(nc, "csem_mixed_solve")
        cont = 0
        for r in range(R):
            c, rel[r] = finish(r, C.c_void_p(dxm.value + r * n * 16), rel[r])
            cont += c
        return nc, cont

    # the same three solves as independent single right-hand-side solves on
    # three host threads, each on its own stream and solve slot
    def mixed_streams_step():
        out = [None] * R

        def run(r):
            itr, rr = np.zeros(1, np.int32), np.zeros(1)
            nc = L.csem_mixed_solve(model, 1, C.byref(fr[r]), C.c_void_p(dxs.value + r * n * 16),
                                    C.c_void_p(db.value + r * n * 16), tol_arr[r:r + 1].ctypes.data, C.byref(opts),
                                    itr.ctypes.data, rr.ctypes.data, None)
            c, relr = finish(r, C.c_void_p(dxs.value + r * n * 16), float(rr[0])) if nc >= 0 else (0, float(rr[0]))
            out[r] = (nc, int(itr[0]), relr, c)

        th = [threading.Thread(target=run, args=(r,)) for r in range(R)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        for o in out:
            _lib.check(o[0], "csem_mixed_solve")
        return out

    # the timed steps run with the kernel timers off: an event pair per
    # launch would add the event packets' time to every launch
    def timed(fn):
        for _ in range(args.warmup):
            fn()
        L.csem_sync(None)
        t0 = time.perf_counter()
        res = None
        for _ in range(args.steps):
            res = fn()
        L.csem_sync(None)
        dt = time.perf_counter() - t0
        return dt, res

    # solves/s counts the right-hand sides that end below their tol
    dtm, (nconv, ncont) = timed(mixed_step)
    okm = int(sum(rel[r] < tols[r] for r in range(R)))
    mixed = {"batch": {"solves_per_s": round(okm * args.steps / dtm, 4), "ms_per_step": round(1e3 * dtm / args.steps, 3),
                       "converged": int(nconv), "fp64_continued": ncont, "inner_iterations": [int(i) for i in its],
                       "true_relres": [float("%.3e" % r) for r in rel]}, "tol": tols}
    dts, sres = timed(mixed_streams_step)
    oks = int(sum(o[2] < tols[r] for r, o in enumerate(sres)))
    mixed["streams"] = {"solves_per_s": round(oks * args.steps / dts, 4), "ms_per_step": round(1e3 * dts / args.steps, 3),
                        "converged": sum(o[0] for o in sres), "fp64_continued": sum(o[3] for o in sres),
                        "inner_iterations": [o[1] for o in sres], "true_relres": [float("%.3e" % o[2]) for o in sres]}
    best = "batch" if mixed["batch"]["solves_per_s"] >= mixed["streams"]["solves_per_s"] else "streams"
    mixed["value_from"] = best
    mixed["solves_per_s"] = mixed[best]["solves_per_s"]
    mixed["ms_per_step"] = mixed[best]["ms_per_step"]
    rows, planes = C.c_int(), C.c_int()
    if L.csem_mixed_tuning(model, R, C.byref(rows), C.byref(planes)) == 0:
        mixed["autotuned_apply_tile"] = {"rows": rows.value, "planes": planes.value}
    plain = None
    if not args.no_plain:
        dtp, pits = timed(plain_step)
        plain = {"solves_per_s": round(R * args.steps / dtp, 4), "ms_per_step": round(1e3 * dtp / args.steps, 3),
                 "iterations": pits}
        # receivers of both solutions (extract_emf on the host, channel-normalised difference)
        xm = np.zeros(R * n, np.complex128)
        xp = np.zeros(R * n, np.complex128)
        L.csem_memcpy_d2h(xm.ctypes.data, dxm if best == "batch" else dxs, R * n * 16, None)
        L.csem_memcpy_d2h(xp.ctypes.data, dxp, R * n * 16, None)
        L.csem_sync(None)
        errs = []
        for r in range(R):
            ses.set_freq(freqs[r])  # the H channels scale by 1 / (i w mu0)
            a = ses.extract(xm[r * n:(r + 1) * n]).astype(np.complex128)
            b = ses.extract(xp[r * n:(r + 1) * n]).astype(np.complex128)
            # per channel (Ex, Ey, Ez, Hx, Hy, Hz), relative to that channel's
            # largest value, and per receiver relative to its total E (and H)
            sc = np.abs(b).max(axis=1)
            e = {"per_channel": [float("%.2e" % (np.abs(a[c] - b[c]).max() / sc[c])) if sc[c] > 0 else 0.0
                                 for c in range(a.shape[0])]}
            for key, cs in (("vs_total_field", slice(0, 3)), ("vs_total_h_field", slice(3, 6))):
                if a[cs].shape[0]:
                    tot = np.sqrt((np.abs(b[cs]) ** 2).sum(axis=0))
                    e[key] = float("%.2e" % (np.abs(a[cs] - b[cs]).max(axis=0) / tot).max())
            errs.append(e)
        mixed["receiver_err_vs_plain_fp64"] = errs
        # true FP64 residual of the plain solutions (the reference stops on its recursive residual)
        tr = []
        dw = C.c_void_p()
        _lib.check(L.csem_malloc(C.byref(dw), n * 16), "csem_malloc")
        for r in range(R):
            L.csem_apply(model, C.byref(fr[r]), C.c_void_p(dxp.value + r * n * 16), dw, None)
            w = np.zeros(n, np.complex128)
            L.csem_memcpy_d2h(w.ctypes.data, dw, n * 16, None)
            L.csem_sync(None)
            br = hb[r * n:(r + 1) * n]
            tr.append(float("%.3e" % (np.linalg.norm(br - w) / np.linalg.norm(br))))
        L.csem_free(dw)
        plain["true_relres"] = tr
    # per-kernel times of the mixed path, in the mode the value comes from:
    # the same solves once more, launched eagerly (graph replays are
    # invisible to the launch timers), one after the other so that no
    # kernel shares the GPU with another solve's, outside the timed region;
    # each launch between its own event pair.  Byte totals count only the
    # right-hand sides a batched launch works on.
    popts = _lib.MixedOpts.default(max_inner=args.niter, use_graph=0)
    L.csem_kernel_timing(1)
    if best == "batch":
        L.csem_mixed_solve(model, R, fr, dxm, db, tol_arr.ctypes.data, C.byref(popts), its.ctypes.data,
                           rel.ctypes.data, None)
    else:
        for r in range(R):
            itr, rr = np.zeros(1, np.int32), np.zeros(1)
            L.csem_mixed_solve(model, 1, C.byref(fr[r]), C.c_void_p(dxs.value + r * n * 16),
                               C.c_void_p(db.value + r * n * 16), tol_arr[r:r + 1].ctypes.data, C.byref(popts),
                               itr.ctypes.data, rr.ctypes.data, None)
    L.csem_sync(None)
    stm = _lib.kernel_stats()
    L.csem_kernel_timing(0)
    # the rooflines: every kernel of one eager solve timed per launch; then
    # the same pass with warm-cache repeat loops, reported beside it
    kl, dominant = roofline_pass(L, model, n, fr[0], db, dxm, dxp)
    kw, _ = roofline_pass(L, model, n, fr[0], db, dxm, dxp, loops=True)
    warm = {k: {x: v[x] for x in ("loop_us", "loop_launches", "loop_gbs")} for k, v in kw.items() if "loop_us" in v}
    copy_gbs = _copy_rate(L)
    dd = ses.dims
    lu_dims = (dd["n1"], dd["n2"], dd["nz"] + 2 * dd["nb"])
    out = {
        "metric": METRIC,
        "value": mixed["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 inner / c128 refinement",
        "data": "synthetic",
        "config": {"workload": "%s: 100x100x100 cells + nb=10 padding, airbc=1 (1024^2 transform), LU-SGS + "
                               "F-BiCGStab, f=%s Hz, tol=1e-8/f, niter %d; value = batched mixed-precision "
                               "solves/s" % (args.workload, "/".join("%g" % f for f in freqs), args.niter),
                   "n_op": n, "solves_per_step": R},
        "mixed": mixed,
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": round(mixed["solves_per_s"] / plain["solves_per_s"], 3) if plain else None,
        "roofline": _roofline(kl, dominant, copy_gbs),
        "roofline_what": ("dominant kernel of the mixed solve by summed time (%s), in the roofline pass: an eager "
                          "single-right-hand-side solve of %s at %g Hz, 40 inner iterations; duration = average "
                          "launch, each between its own HIP-event pair on the solve's stream (the command "
                          "bench.py --roofline-only runs the same pass alone)" % (dominant, args.workload, freqs[0])),
        "apply_roofline": _roofline(kl, "k_mx_apply", copy_gbs),
        "apply_fp64_roofline": _roofline(kl, "k_apply", copy_gbs),
        "hbm_copy": {"achieved": round(copy_gbs, 1), "unit": "GB/s", "frac": round(copy_gbs / HBM_PEAK_GBS, 4),
                     "what": "k_copy16, 1 GiB device copy, bytes read + written / time"},
        "lusgs_roofline_mixed": _lusgs_roofline(kl, lu_dims, True),
        "lusgs_roofline_fp64": _lusgs_roofline(kl, lu_dims, False),
        "kernels": kl,
        "kernels_warm_loops": warm,
        "kernels_warm_loops_what": ("the same pass with each of the first %d launches of a kernel followed by %d "
                                    "back-to-back launches with the same arguments between one event pair "
                                    "(csem_kernel_repeat): the kernel with its inputs warm in the caches"
                                    % (LOOP_LOOPS, LOOP_REPS)),
        "kernel_ms_mixed": {k: round(v[0], 3) for k, v in sorted(stm.items(), key=lambda kv: -kv[1][0])},
        "kernel_bytes_per_launch_mixed": {k: int(v[2] / v[1]) for k, v in stm.items() if v[1]},
        "kernel_ms_mixed_what": "summed kernel time of one step's %d solves (%s mode), launched eagerly one "
                                "solve after the other, each launch between its own event pair (which adds the "
                                "event packets' few microseconds to every launch)" % (R, best),
    }
    if args.serve_requests > 0:
        # serving: a warm engine (models resident) fed seeded requests on
        # 64^3 and 100^3 grids at Poisson arrival times
        from demo_matrixfree_csem_amd import serve
        reqs = serve.quick_requests(args.serve_requests, 4)
        eng = serve.Engine(workers=3, rhs_per_solve=1,
                           opts=_lib.MixedOpts.default(use_graph=0) if args.no_graph else None)
        try:
            eng.preload(reqs)
            _, sst = eng.run(reqs, serve.poisson_arrivals(len(reqs), args.serve_rate, 4))
        finally:
            eng.close()
        sst["rate_per_s"] = args.serve_rate
        sst["stream"] = "64^3 and 100^3 grids (two models each), 1-3 frequencies, 1-2 sources per request"
        out["serving"] = sst
    if args.serve_cold > 0:
        # cold serving over the full grid set (64^3, 100^3, 128x128x96,
        # 180x120x120): the capacity first (a warm engine, every request at
        # t = 0), then a fresh engine with no model built, fed the same
        # requests at Poisson arrivals of 0.7 times that capacity: its
        # p50/p99 include the model builds (host setup and upload)
        from demo_matrixfree_csem_amd import serve
        creqs = [surveys.serve_request(7000 + i, grids=[surveys.SERVE_GRIDS[i % len(surveys.SERVE_GRIDS)]],
                                       max_freqs=2, max_sources=1, n_models=2) for i in range(args.serve_cold)]
        eng = serve.Engine(workers=3, rhs_per_solve=1)
        try:
            eng.preload(creqs)
            _, cap = eng.run(creqs, [0.0] * len(creqs))
        finally:
            eng.close()
        cap_rps = len(creqs) / cap["wall_s"]
        rate = 0.7 * cap_rps
        eng = serve.Engine(workers=3, rhs_per_solve=1)
        try:
            _, cst = eng.run(creqs, serve.poisson_arrivals(len(creqs), rate, 7))
        finally:
            eng.close()
        cst.update({"rate_per_s": round(rate, 4), "capacity_requests_per_s": round(cap_rps, 4),
                    "capacity_solves_per_s": cap["solves_per_s"], "warm": False,
                    "grids": sorted({"%dx%dx%d" % r[0][0] for r in creqs}),
                    "stream": "the four grids in turn, one of two seeded models each, 1-2 frequencies, one source; "
                              "arrivals at 0.7 of the capacity measured with the models resident"})
        out["serving_cold"] = cst
    if not args.no_cpu and plain:
        out["cpu_baseline"] = cpu_baseline(survey, freqs, args.cpu_iters, plain["iterations"])
    ses.close()
    if not args.no_large:
        out["apply_large"] = apply_large(copy_gbs=copy_gbs)
    out.update(legs())
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
