"""Host side of the serving loop (no GPU): the request generator, the
arrival process, and the per-request source / receiver helpers, which must
give what the reference's own inject_source / extract_emf give for a survey
whose acquisition files hold that source and those receivers
(inject_extract.c:24-130, 132-200)."""
import os

import numpy as np

from demo_matrixfree_csem_amd import surveys
from demo_matrixfree_csem_amd.serve import poisson_arrivals, solve_tol
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
