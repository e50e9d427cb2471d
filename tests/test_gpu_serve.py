"""Serving loop on the GPU: every request's receivers are the same served
alone and served in a mix of overlapping requests on several models, and
admission keeps the charged device footprint within the budget, evicting
idle models when it must.

Tolerance: receivers within 1e-6 of the channel's largest value (the mixed
solve's receiver tolerance against the FP64 oracle); each solve's true
FP64 relative residual below its tol = 1e-8/f.
"""
import numpy as np
import pytest

from demo_matrixfree_csem_amd import _lib, surveys
from demo_matrixfree_csem_amd.serve import Engine, solve_tol

pytestmark = pytest.mark.gpu
SMALL = [(24, 24, 16), (32, 28, 20)]


def _requests(n, seed, n_models=2):
    return [surveys.serve_request(seed * 100 + i, grids=SMALL, max_freqs=2, max_sources=2, nrec=(9, 17),
                                  n_models=n_models, nfft=64) for i in range(n)]


def _rec_err(a, b):
    a, b = a.astype(np.complex128), b.astype(np.complex128)
    scale = np.abs(b).max(axis=1, keepdims=True)
    return float((np.abs(a - b) / np.where(scale > 0, scale, 1)).max())


def _check_converged(res):
    for (s, f), (rec, it, rr) in res["receivers"].items():
        assert rr < solve_tol(f), (s, f, it, rr)


def test_each_request_same_alone_and_in_the_mix():
    reqs = _requests(6, 1)
    print("case serve mix: %d requests on %d models" % (len(reqs), len({r[0] for r in reqs})), flush=True)
    mix = Engine(workers=3, rhs_per_solve=2)
    try:
        res_mix, st = mix.run(reqs, [0.01 * i for i in range(len(reqs))])
    finally:
        mix.close()
    print(st, flush=True)
    assert st["solves"] == sum(len(r[2]) * len(r[3]) for r in reqs)
    assert st["peak_admitted_bytes"] <= st["budget_bytes"]
    alone = Engine(workers=1, rhs_per_solve=1)
    try:
        for i, req in enumerate(reqs):
            print("case serve alone: request %d" % i, flush=True)
            res, _ = alone.run([req], [0.0])
            _check_converged(res_mix[i])
            _check_converged(res[0])
            assert set(res[0]["receivers"]) == set(res_mix[i]["receivers"])
            for k, (rec, it, rr) in res[0]["receivers"].items():
                e = _rec_err(res_mix[i]["receivers"][k][0], rec)
                assert e < 1e-6, (i, k, e)
    finally:
        alone.close()


def test_admission_evicts_idle_models_within_budget():
    reqs = _requests(6, 2, n_models=4)
    keys = {r[0] for r in reqs}
    probe = Engine(workers=1)
    try:
        probe.preload(reqs)
        # one model resident with two solves of two right-hand sides running
        # (two pooled slots, two more charged) and both workers' buffers
        one = max(m.bytes() + 4 * m.slot_bytes(2) + 2 * 2 * (2 * m.n * 16) for m in probe.models.values())
    finally:
        probe.close()
    budget = int(1.1 * one)
    print("case serve eviction: %d models, budget %d bytes" % (len(keys), budget), flush=True)
    eng = Engine(budget_bytes=budget, workers=2, rhs_per_solve=2)
    try:
        res, st = eng.run(reqs, [0.0] * len(reqs))
        print(st, flush=True)
        assert st["peak_admitted_bytes"] <= budget
        assert all(r is not None for r in res)
        for r in res:
            _check_converged(r)
        if len(keys) > 2:
            assert st["evictions"] >= 1
    finally:
        eng.close()


def test_engine_finishes_a_stalled_solve_in_fp64(workdir):
    """A request the mixed solve does not finish within its refinement
    budget (the anisotropic model at 0.25 Hz, strongly stretched padding,
    one refinement step allowed): the engine continues it with the FP64
    solver and serves receivers equal to a plain FP64 solve."""
    from demo_matrixfree_csem_amd.session import Session
    s = surveys.vti_marine(2, 36, 24, 24, 33)
    src = [tuple(float(v) for v in s.acq.sources[0][:3])]
    recs = [tuple(float(v) for v in r[:3]) for r in s.acq.receivers]
    req = (("vti_marine", 36, 24, 24), lambda: s, [0.25], src, recs)
    print("case serve: stalled mixed solve finished in FP64", flush=True)
    eng = Engine(workers=1, rhs_per_solve=1, opts=_lib.MixedOpts.default(max_outer=1))
    try:
        res, st = eng.run([req], [0.0])
    finally:
        eng.close()
    print(st, flush=True)
    assert st["fp64_continued"] == 1
    _check_converged(res[0])
    rec, it, rr = res[0]["receivers"][(0, 0.25)]
    surveys.write_files(workdir, s)
    ses = Session(s.argv(workdir))
    try:
        ses.set_freq(0.25)
        b = ses.inject_at(*src[0])
        x, itp = ses.solve(b, solve_tol(0.25), niter=4000, algopt=1)
        assert 0 < itp < 4000
        want = ses.extract_at(x, recs)
    finally:
        ses.close()
    e = _rec_err(rec, want)
    print("served after %d iterations (relres %.2e); FP64 alone %d; receiver error %.2e" % (it, rr, itp, e), flush=True)
    assert e < 1e-6, e


def test_request_file_served_and_written(workdir):
    """A request file (serve.REQUEST_SCHEMA) served with two right-hand
    sides per solve on one worker: the two single-pair requests on the
    file-given model share one solve (cross-request batching); every
    request's emf_0001.txt is in the reference's format and its receivers
    equal those of this package's csem_main (the reference's FP64 solver)
    on the same inputs within 1e-6 of the channel's largest value, and the
    double-precision file holds what the float file rounds."""
    import json
    import os
    import subprocess
    from demo_matrixfree_csem_amd.serve import load_requests, write_results
    s = surveys.tiny(6, nx=16, ny=14, nz=10)
    s.flags.update(n1fft=64, n2fft=64, niter=1000)
    mdir = os.path.join(workdir, "model")
    surveys.write_files(mdir, s)
    files = {k: os.path.join("model", k) for k in ("fx1", "fx2", "fx3", "frho11", "frho22", "frho33")}
    src = [float(v) for v in s.acq.sources[0][:3]]
    recs = [[float(v) for v in r[:3]] for r in s.acq.receivers]
    reqs = [{"id": "p%d" % k, "model": {"files": files, "flags": {"n1fft": 64, "nb": s.flags["nb"]}},
             "freqs": [f], "sources": [src], "receivers": recs} for k, f in enumerate((0.5, 1.25))]
    path = os.path.join(workdir, "reqs.jsonl")
    with open(path, "w") as fh:
        fh.write("\n".join(json.dumps(r) for r in reqs) + "\n")
    loaded = load_requests(path)
    print("case serve request file", flush=True)
    eng = Engine(workers=1, rhs_per_solve=2)
    try:
        res, st = eng.run([r[2] for r in loaded], [r[1] for r in loaded])
    finally:
        eng.close()
    print(st, flush=True)
    assert st["batches"] == 1 and st["cross_request_batches"] == 1
    out = os.path.join(workdir, "out")
    write_results(out, [r[0] for r in loaded], [r[2] for r in loaded], res)
    # the reference's solver on the same inputs, one frequency per run
    main = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "demo_matrixfree_csem_amd",
                        "csem_main")
    for k, f in enumerate((0.5, 1.25)):
        _check_converged(res[k])
        run = os.path.join(workdir, "main%d" % k)
        os.makedirs(run)
        s.freqs = [f]
        r = subprocess.run([main] + s.argv(mdir), cwd=run, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        ref = np.loadtxt(os.path.join(run, "emf_0001.txt"), skiprows=1)
        got = np.loadtxt(os.path.join(out, "p%d" % k, "emf_0001.txt"), skiprows=1)
        g64 = np.loadtxt(os.path.join(out, "p%d" % k, "emf_0001_f64.txt"), skiprows=1)
        assert np.array_equal(got[:, :4], ref[:, :4])
        v, w = got[:, 4] + 1j * got[:, 5], ref[:, 4] + 1j * ref[:, 5]
        v64 = g64[:, 4] + 1j * g64[:, 5]
        for ch in range(3):
            sel = ref[:, 2] == ch + 1
            scale = np.abs(w[sel]).max()
            e = np.abs(v[sel] - w[sel]).max() / scale
            print("request p%d channel %d: served vs csem_main %.2e" % (k, ch, e), flush=True)
            assert e < 1e-6, (k, ch, e)
        # the float table is the double one rounded, to its printed digits
        for a, b in zip(open(os.path.join(out, "p%d" % k, "emf_0001.txt")).read().splitlines()[1:], v64):
            t = a.split()
            c = np.complex64(b)
            assert t[4:] == ["%e" % c.real, "%e" % c.imag]
