"""Host side of the serving loop (no GPU): the request generator, the
arrival process, and the per-request source / receiver helpers, which must
give what the reference's own inject_source / extract_emf give for a survey
whose acquisition files hold that source and those receivers
(inject_extract.c:24-130, 132-200); the request-file schema and the
result files (write_data.c:24-34)."""
import json
import os

import numpy as np
import pytest

from demo_matrixfree_csem_amd import surveys
from demo_matrixfree_csem_amd.serve import load_requests, poisson_arrivals, solve_tol, write_results
from demo_matrixfree_csem_amd.session import Session

SMALL = [(24, 24, 16), (32, 28, 20)]


def _session(workdir, survey, name):
    d = os.path.join(workdir, name)
    surveys.write_files(d, survey)
    return Session(survey.argv(d), device=False)


def _request_with_sources(min_sources=2):
    for seed in range(100):
        r = surveys.serve_request(seed, grids=SMALL, nfft=64)
        if len(r[3]) >= min_sources:
            return r
    raise AssertionError("no request with %d sources" % min_sources)


def test_serve_request_is_seeded_and_in_range():
    a = surveys.serve_request(7)
    b = surveys.serve_request(7)
    assert a[0] == b[0] and a[2] == b[2] and a[3] == b[3] and a[4] == b[4]
    for seed in range(40):
        key, build, freqs, srcs, recs = surveys.serve_request(seed)
        assert key[0] in surveys.SERVE_GRIDS
        assert 1 <= len(freqs) <= 6 and all(0.125 <= f <= 1.25 for f in freqs)
        assert freqs == sorted(set(freqs))
        assert 1 <= len(srcs) <= 4 and 33 <= len(recs) <= 121
        assert all(s[2] == 950.0 for s in srcs) and all(r[2] == 1000.0 for r in recs)


def test_serve_models_are_shared_between_requests():
    keys = [surveys.serve_request(seed)[0] for seed in range(60)]
    assert len(set(keys)) < len(keys)


def test_poisson_arrivals():
    t = poisson_arrivals(2000, 4.0, 1)
    assert all(b > a for a, b in zip(t, t[1:]))
    gaps = np.diff([0.0] + t)
    assert abs(gaps.mean() - 0.25) < 0.02


def test_solve_tol_matches_the_driver():
    # main.c forms tol/f in single precision
    assert solve_tol(0.25) == float(np.float32(np.float32(1e-8) / np.float32(0.25)))


def test_inject_at_equals_reference_inject(workdir):
    key, build, freqs, srcs, recs = _request_with_sources()
    s = build()
    ses = _session(workdir, s, "a")
    ses.set_freq(freqs[0])
    b0 = ses.inject()
    assert np.array_equal(ses.inject_at(*srcs[0]), b0)
    # another source: the reference's inject for a survey whose source file holds it
    s2 = build()
    s2.acq.sources = surveys._acq_table(srcs[-1:])
    ses2 = _session(workdir, s2, "b")
    ses2.set_freq(freqs[0])
    b = ses.inject_at(*srcs[-1])
    assert np.array_equal(b, ses2.inject())
    assert not np.array_equal(b, b0)
    # the session's own source is restored
    assert np.array_equal(ses.inject(), b0)


def test_extract_at_equals_reference_extract(workdir):
    key, build, freqs, srcs, recs = _request_with_sources(1)
    ses = _session(workdir, build(), "a")
    rng = np.random.default_rng(0)
    x = rng.standard_normal(ses.n) + 1j * rng.standard_normal(ses.n)
    full = ses.extract(x)
    assert np.array_equal(ses.extract_at(x, recs), full)
    assert np.array_equal(ses.extract_at(x, recs[3:7]), full[:, 3:7])
    # the session's receivers are restored
    assert np.array_equal(ses.extract(x), full)


def _request_file(workdir):
    """A request file with a model given by the reference's binary inputs
    (written by surveys.write_files) and one synthetic model."""
    s = surveys.tiny(5, nx=12, ny=10, nz=8)
    mdir = os.path.join(workdir, "model")
    surveys.write_files(mdir, s)
    files = {k: os.path.join("model", k) for k in ("fx1", "fx2", "fx3", "frho11", "frho22", "frho33")}
    reqs = [
        {"id": "a", "arrival_s": 0.5, "model": {"files": files, "flags": {"n1fft": 64}},
         "freqs": [0.75, 0.25], "sources": [[0.0, 0.0, 950.0]], "receivers": [[-200.0, 0.0, 1000.0], [200.0, 0.0, 1000.0]]},
        {"id": "b", "model": {"files": files, "flags": {"n1fft": 64}},
         "freqs": [0.5], "sources": [[10.0, 0.0, 950.0], [-10.0, 0.0, 950.0]], "receivers": [[0.0, 0.0, 1000.0]]},
        {"id": "c", "model": {"synthetic": {"grid": [24, 24, 16], "kind": "grf", "seed": 1003, "median": 4.0},
                              "flags": {"n1fft": 64, "chrec": "Ex,Ez"}},
         "freqs": [1.25], "sources": [[0.0, 0.0, 950.0]], "receivers": [[0.0, 0.0, 1000.0]]},
    ]
    path = os.path.join(workdir, "reqs.jsonl")
    with open(path, "w") as f:
        for r in reqs:
            f.write(json.dumps(r) + "\n")
    return s, path


def test_request_file_schema(workdir):
    """load_requests: models from the reference's binary inputs are read
    back exactly and shared between requests with the same inputs and
    flags; synthetic models are the survey generator's; bad inputs are
    refused."""
    s, path = _request_file(workdir)
    loaded = load_requests(path)
    assert [r[0] for r in loaded] == ["a", "b", "c"] and [r[1] for r in loaded] == [0.5, 0.0, 0.0]
    (ka, ba, fa, sa, ra), (kb, bb, *_), (kc, bc, fc, *_) = [r[2] for r in loaded]
    assert ka == kb and ka != kc
    assert fa == [0.25, 0.75] and sa == [(0.0, 0.0, 950.0)] and len(ra) == 2
    sv = ba()
    for k in ("x1node", "x2node", "x3node", "rho11", "rho22", "rho33"):
        assert np.array_equal(getattr(sv.model, k), np.asarray(getattr(s.model, k), np.float32)), k
    assert sv.flags["n1fft"] == 64 and "lextend" not in sv.flags
    sc = bc()
    want = surveys.marine(1003, 24, 24, 16, kind="grf", median=4.0)
    assert np.array_equal(sc.model.rho11, want.rho11) and sc.flags["chrec"] == "Ex,Ez"
    # a resistivity file of the wrong size is refused
    bad = os.path.join(workdir, "model", "frho22")
    np.zeros(5, np.float32).tofile(bad)
    with pytest.raises(ValueError):
        load_requests(path)[0][2][1]()
    with open(path, "a") as f:
        f.write(json.dumps({"id": "a", "model": {"synthetic": {"grid": [8, 8, 8]}}, "freqs": [1.0],
                            "sources": [[0, 0, 0]], "receivers": [[0, 0, 0]]}) + "\n")
    with pytest.raises(ValueError):
        load_requests(path)


def test_write_results_format(workdir):
    """write_results: the reference's table (write_data.c:24-34: channel,
    frequency, receiver; float32 printed with %e) per source, the double
    values beside it, and a result.json."""
    freqs, srcs, recs = [0.25, 0.75], [(0, 0, 950), (5, 0, 950)], [(0, 0, 1000), (1, 0, 1000), (2, 0, 1000)]
    rng = np.random.default_rng(3)
    res = {"latency": 1.25, "receivers": {}}
    for s in range(2):
        for f in freqs:
            r = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
            res["receivers"][(s, f)] = (r * 1e-9, 100 + s, 1e-9)
    write_results(workdir, ["q"], [(None, None, freqs, srcs, recs)], [res])
    d = os.path.join(workdir, "q")
    for s in range(2):
        rows = open(os.path.join(d, "emf_%04d.txt" % (s + 1))).read().splitlines()
        rows64 = open(os.path.join(d, "emf_%04d_f64.txt" % (s + 1))).read().splitlines()
        assert rows[0].split() == ["iTx", "iRx", "ichrec", "ifreq", "emf_real", "emf_imag"]
        assert len(rows) == 1 + 3 * 2 * 3 and len(rows64) == len(rows)
        k = 1
        for ch in range(3):
            for fi, f in enumerate(freqs):
                for ir in range(3):
                    t, t64 = rows[k].split(), rows64[k].split()
                    v = res["receivers"][(s, f)][0][ch, ir]
                    assert [int(x) for x in t[:4]] == [s + 1, ir + 1, ch + 1, fi + 1]
                    v32 = np.complex64(v)
                    assert t[4:] == ["%e" % v32.real, "%e" % v32.imag]
                    assert complex(float(t64[4]), float(t64[5])) == v
                    k += 1
    info = json.load(open(os.path.join(d, "result.json")))
    assert info["latency_s"] == 1.25 and len(info["solves"]) == 4
