#!/usr/bin/env python3
"""Converged solves per second of the single-rank frequency-domain CSEM
solve on one MI355X.

Workload (one step): the 100^3-cell layered marine model (100 m cells, 8
seeded log-uniform layers under 1000 m of 0.3125 ohm-m seawater, nb=10,
airwave boundary with a 1024^2 transform), one Ex source, three
frequencies 0.25 / 0.75 / 1.25 Hz solved to tol = 1e-8/f with LU-SGS +
F-BiCGStab in complex FP64 -- the reference's path, device-resident.
Inputs (model, right-hand sides) are resident in HBM before the timed
region; every step solves all three frequencies from x = 0.

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
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/s @ tol 1e-8: batched mixed vs plain FP64 GPU; apply HBM TB/s; p99 lat"
HBM_PEAK_GBS = 8000.0


def _roofline(stats, name):
    ms, cnt, by = stats[name]
    avg_s = ms / cnt / 1e3
    per_launch = by / cnt
    gbs = per_launch / avg_s / 1e9
    return {"kernel": name, "bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
            "bytes_per_launch": int(per_launch), "avg_us": round(avg_s * 1e6, 3), "launches": cnt}


def cpu_baseline(survey, freq, iters, gpu_iters):
    """The oracle (serial C restatement of the reference path) on this
    host: `iters` LU-SGS F-BiCGStab iterations of one solve, scaled to
    solves/s with the iteration counts the GPU needed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    m = survey.model
    f = survey.flags
    orc = O.Oracle(m.x1node, m.x2node, m.x3node, m.rho11, m.rho22, m.rho33, nb=f["nb"], airbc=f["airbc"],
                   fvm=f.get("fvm", 0), nfft=1024)
    orc.set_freq(freq)
    b = orc.inject(0, *survey.acq.sources[0][:3])
    t0 = time.perf_counter()
    orc.fbicgstab(b, iters, 0.0)
    dt = time.perf_counter() - t0
    per_iter = dt / iters
    mean_iters = float(np.mean([i + 1 for i in gpu_iters]))
    return {"value": round(1.0 / (per_iter * mean_iters), 6), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": "%d LU-SGS F-BiCGStab iterations of the %.2f Hz solve on the same model (serial C "
                      "restatement, 1 thread, %.1f s, %.3f s/iteration), scaled by the GPU's mean of %.1f "
                      "iterations per solve" % (iters, freq, dt, per_iter, mean_iters)}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="layered_100")
    ap.add_argument("--cpu-iters", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--niter", type=int, default=1000)
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
    tol0 = np.float32(survey.flags.get("tol", 1e-8))

    # resident inputs: one right-hand side per frequency, and x
    rhs = []
    for f in freqs:
        ses.set_freq(f)
        b = ses.inject()
        d = C.c_void_p()
        _lib.check(L.csem_malloc(C.byref(d), n * 16), "csem_malloc")
        _lib.check(L.csem_memcpy_h2d(d, b.ctypes.data, n * 16, None), "upload")
        rhs.append((f, freq_consts(f), float(np.float32(tol0 / np.float32(f))), d))
    x = C.c_void_p()
    _lib.check(L.csem_malloc(C.byref(x), n * 16), "csem_malloc")
    hist = np.zeros(args.niter + 2)

    def step():
        its = []
        for f, fc, tol, d in rhs:
            L.csem_memset(x, 0, n * 16, None)
            it = L.csem_bicgstab(model, C.byref(fc), x, d, args.niter, tol, 1, hist.ctypes.data, None)
            _lib.check(it, "csem_bicgstab")
            its.append(it)
        return its

    for _ in range(args.warmup):
        step()
    L.csem_sync(None)
    L.csem_kernel_timing(1)
    t0 = time.perf_counter()
    iters = []
    for _ in range(args.steps):
        iters = step()
    L.csem_sync(None)
    dt = time.perf_counter() - t0
    stats = _lib.kernel_stats()
    L.csem_kernel_timing(0)

    dominant = max(stats, key=lambda k: stats[k][0])
    nsolves = args.steps * len(rhs)
    out = {
        "metric": METRIC,
        "value": round(nsolves / dt, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {"workload": "%s: 100x100x100 cells + nb=10 padding, airbc=1 (1024^2 transform), LU-SGS + "
                               "F-BiCGStab, f=%s Hz, tol=1e-8/f, plain FP64 device-resident path"
                               % (args.workload, "/".join("%g" % f for f in freqs)),
                   "n_op": n, "solves_per_step": len(rhs), "iterations": iters},
        "roofline": _roofline(stats, dominant),
        "apply_roofline": _roofline(stats, "k_apply") if "k_apply" in stats else None,
        "kernel_ms": {k: round(v[0] / args.steps, 3) for k, v in sorted(stats.items(), key=lambda kv: -kv[1][0])},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(survey, freqs[0], args.cpu_iters, iters)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
