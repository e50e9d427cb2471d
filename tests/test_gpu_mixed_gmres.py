"""Mixed-precision solve with the FGMRES inner method (csem_mixed_opts.krylov
= KRYLOV_FGMRES): FP32 FGMRES(m) cycles with LU-SGS inside FP64 refinement,
checked against the FP64 oracle's FGMRES and F-BiCGStab (the pinned CPU
restatement of the reference, solver.c:93-178 and 403-512).

Tolerances (stated here and in the assertions):
  * every right-hand side reaches its tol = 1e-8/f in the true FP64
    residual |b - A x| / |b| computed by the oracle's operator;
  * every receiver value within 1e-6 of the channel's largest value of the
    oracle's FP64 FGMRES(10) solution (and of its F-BiCGStab solution):
    both solvers stop below the same tol, so their receivers differ by
    the tolerance, not by FP32 rounding;
  * a right-hand side solved alone and in a batch of three: bit-identical
    (the FP32 kernels and every per-RHS reduction do not depend on R);
  * graph replay of a cycle and eager launches: bit-identical;
  * restart lengths 4, 7 and 16 (the three compiled basis sizes of the
    orthogonalisation passes) converge to the same bound.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import oracle as O  # noqa: E402

from demo_matrixfree_csem_amd import surveys, _lib  # noqa: E402
from demo_matrixfree_csem_amd.session import Session  # noqa: E402

pytestmark = pytest.mark.gpu
FREQS = [0.25, 0.75, 1.25]


def _setup(workdir, seed=2):
    s = surveys.tiny(seed, airbc=1)
    s.flags.update(n1fft=64, n2fft=64, chrec="Ex,Ey,Ez,Hx,Hy,Hz")
    s.freqs = list(FREQS)
    surveys.write_files(workdir, s)
    ses = Session(s.argv(workdir))
    m = s.model
    orc = O.Oracle(m.x1node, m.x2node, m.x3node, m.rho11, m.rho22, m.rho33, nb=s.flags["nb"], airbc=1, nfft=64)
    return s, ses, orc


def _tol(f):
    return float(np.float32(np.float32(1e-8) / np.float32(f)))


def _receivers(ses, x, f):
    ses.set_freq(f)
    return ses.extract(x).astype(np.complex128)


def _rec_err(a, b):
    """max over channels of max |a - b| / max |b| (channel-normalised)."""
    scale = np.abs(b).max(axis=1, keepdims=True)
    return float((np.abs(a - b) / np.where(scale > 0, scale, 1)).max())


def _gm(**kw):
    kw.setdefault("max_inner", 4000)
    return _lib.MixedOpts.default(krylov=_lib.KRYLOV_FGMRES, **kw)


def _rhs(ses, freqs):
    bs, tols = [], []
    for f in freqs:
        ses.set_freq(f)
        bs.append(ses.inject())
        tols.append(_tol(f))
    return bs, tols


def test_mixed_fgmres_matches_fp64_oracle(workdir):
    print("case mixed FGMRES vs oracle", flush=True)
    s, ses, orc = _setup(workdir)
    try:
        bs, tols = _rhs(ses, FREQS)
        xs, its, rel, nconv = ses.solve_mixed(bs, FREQS, tols, _gm())
        assert nconv == 3, (its, rel)
        for r, f in enumerate(FREQS):
            orc.set_freq(f)
            orc.set_red(O.RED_CHUNKED)
            xg, itg, _ = orc.fgmres(bs[r], 400, tols[r], 10)
            xb, itb, _ = orc.fbicgstab(bs[r], 2000, tols[r])
            res = np.linalg.norm(bs[r] - orc.apply(xs[r])) / np.linalg.norm(bs[r])
            assert res < tols[r], (f, res)
            mine = _receivers(ses, xs[r], f)
            eg = _rec_err(mine, _receivers(ses, xg, f))
            eb = _rec_err(mine, _receivers(ses, xb, f))
            print("f=%g FP32 FGMRES steps %d (FP64 FGMRES %d, F-BiCGStab %d) true relres %.2e receiver "
                  "error vs FGMRES %.2e vs F-BiCGStab %.2e" % (f, its[r], itg, itb, res, eg, eb), flush=True)
            assert eg < 1e-6 and eb < 1e-6, (f, eg, eb)
    finally:
        ses.close()


def test_mixed_fgmres_batch_equals_alone(workdir):
    print("case mixed FGMRES batch vs alone", flush=True)
    s, ses, orc = _setup(workdir, seed=3)
    try:
        bs, tols = _rhs(ses, FREQS)
        xs, its, rel, nconv = ses.solve_mixed(bs, FREQS, tols, _gm())
        assert nconv == 3
        for r, f in enumerate(FREQS):
            x1, it1, rel1, nc1 = ses.solve_mixed([bs[r]], [f], [tols[r]], _gm())
            print("f=%g batched %d alone %d steps" % (f, its[r], it1[0]), flush=True)
            assert nc1 == 1 and it1[0] == its[r]
            assert np.array_equal(x1[0], xs[r])
    finally:
        ses.close()


def test_mixed_fgmres_graph_equals_eager(workdir):
    print("case mixed FGMRES graph vs eager", flush=True)
    s, ses, orc = _setup(workdir, seed=4)
    try:
        bs, tols = _rhs(ses, FREQS[:2])
        xg, itg, _, ng = ses.solve_mixed(bs, FREQS[:2], tols, _gm(use_graph=1))
        xe, ite, _, ne = ses.solve_mixed(bs, FREQS[:2], tols, _gm(use_graph=0))
        assert ng == ne == 2
        assert list(itg) == list(ite)
        for a, b in zip(xg, xe):
            assert np.array_equal(a, b)
    finally:
        ses.close()


@pytest.mark.parametrize("m", [4, 7, 16])
def test_mixed_fgmres_restart_lengths(workdir, m):
    print("case mixed FGMRES restart %d" % m, flush=True)
    s, ses, orc = _setup(workdir, seed=5)
    try:
        f = 0.75
        bs, tols = _rhs(ses, [f])
        xs, its, rel, nconv = ses.solve_mixed(bs, [f], tols, _gm(restart=m))
        assert nconv == 1, (its, rel)
        orc.set_freq(f)
        res = np.linalg.norm(bs[0] - orc.apply(xs[0])) / np.linalg.norm(bs[0])
        assert res < tols[0]
        orc.set_red(O.RED_CHUNKED)
        xo, ito, _ = orc.fbicgstab(bs[0], 2000, tols[0])
        e = _rec_err(_receivers(ses, xs[0], f), _receivers(ses, xo, f))
        print("restart %d: %d steps, true relres %.2e, receiver error %.2e" % (m, its[0], res, e), flush=True)
        assert e < 1e-6
    finally:
        ses.close()


def test_mixed_fgmres_rejects_bad_restart(workdir):
    s, ses, orc = _setup(workdir, seed=6)
    try:
        bs, tols = _rhs(ses, [0.75])
        with pytest.raises(RuntimeError):
            ses.solve_mixed(bs, [0.75], tols, _gm(restart=17))
    finally:
        ses.close()
