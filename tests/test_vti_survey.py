"""The anisotropic marine workload (180 x 120 x 120 cells, six frequencies,
BiCGStab and GMRES, six channels): its generator follows the stated
construction, and (on the GPU, at a reduced size) the mixed-precision solve
agrees with both FP64 Krylov drivers on every channel."""
import numpy as np
import pytest

from demo_matrixfree_csem_amd import surveys


def test_vti_marine_construction():
    s = surveys.vti_marine()
    m = s.model
    assert m.shape == (180, 120, 120)
    for ax, centred in ((m.x1node, True), (m.x2node, True), (m.x3node, False)):
        d = np.diff(ax.astype(np.float64))
        assert d.min() >= 50.0 - 1e-3 and d.max() <= 250.0 + 1e-3
        if centred:
            assert abs(float(ax[0]) + float(ax[-1])) < 1e-2
        else:
            assert ax[0] == 0.0
    zc = 0.5 * (m.x3node[1:] + m.x3node[:-1])
    water, below = zc < 1000.0, zc > 1000.0
    assert np.all(m.rho11[water] == surveys.RHO_WATER) and np.all(m.rho33[water] == surveys.RHO_WATER)
    assert np.array_equal(m.rho11, m.rho22)
    assert np.array_equal(m.rho33[below], 2 * m.rho11[below])  # VTI
    r = m.rho11[below]
    res = r >= 50.0
    assert 0.001 < res.mean() < 0.2 and r[res].max() <= 300.0  # three resistive ellipsoids
    assert len(np.unique(r[res])) == 3
    assert abs(np.median(r[~res]) / 2.0 - 1) < 0.1  # log-normal field, median 2 ohm-m
    assert s.freqs == [0.125, 0.25, 0.5, 0.75, 1.0, 1.25]
    assert len(s.acq.receivers) == 121
    f = s.flags
    assert (f["nb"], f["lextend"], f["nrestart"], f["chrec"]) == (10, 120e3, 10, "Ex,Ey,Ez,Hx,Hy,Hz")


def _tol(f):
    return float(np.float32(np.float32(1e-8) / np.float32(f)))


def _rec_err(a, b):
    scale = np.abs(b).max(axis=1, keepdims=True)
    return float((np.abs(a - b) / np.where(scale > 0, scale, 1)).max())


@pytest.mark.gpu
def test_vti_mixed_matches_fp64_bicgstab_and_gmres(workdir):
    """A 36 x 24 x 24 version of the workload (same construction, padding
    to 120 km: strongly stretched cells), frequencies 0.25, 0.75, 1.25 Hz
    batched in one mixed solve.  At 0.25 Hz this model is ill conditioned
    enough that an FP32 inner solve stagnates; the refinement ends such an
    inner solve and restarts from the true residual, and all three
    frequencies converge.  Each matches FP64 F-BiCGStab, and FP64
    FGMRES(10) where that converges (it stagnates at 0.25 Hz), on all six
    channels within 1e-6 of the channel maximum.  A mixed solve cut short
    (one refinement step) is finished by the FP64 solver continued from its
    solution, with the same receivers as from zero."""
    import ctypes as C
    import threading
    from demo_matrixfree_csem_amd import _lib, serve
    from demo_matrixfree_csem_amd.session import Session
    s = surveys.vti_marine(2, 36, 24, 24, 33)
    surveys.write_files(workdir, s)
    freqs = [0.25, 0.75, 1.25]
    print("case vti_marine 36x24x24 mixed vs fp64 bicgstab and gmres", flush=True)
    ses = Session(s.argv(workdir))
    try:
        bs, tols = [], []
        for f in freqs:
            ses.set_freq(f)
            bs.append(ses.inject())
            tols.append(_tol(f))
        xs, its, rel, nconv = ses.solve_mixed(bs, freqs, tols, _lib.MixedOpts.default(max_inner=6000))
        print("mixed: iterations %s, relres %s, converged %d" % (list(its), list(rel), nconv), flush=True)
        assert nconv == 3 and all(r < t for r, t in zip(rel, tols))
        ngmres, rbs = 0, []
        for r, f in enumerate(freqs):
            ses.set_freq(f)
            xb, itb = ses.solve(bs[r], tols[r], niter=6000, algopt=1)
            assert 0 < itb < 6000, (f, itb)
            rb = ses.extract(xb).astype(np.complex128)
            rbs.append(rb)
            assert rb.shape == (6, 33) and np.all(np.abs(rb[4]) > 0)
            res = np.linalg.norm(bs[r] - ses.fdfd_apply(xs[r])) / np.linalg.norm(bs[r])
            assert res < 2 * tols[r], (f, res)
            rm = ses.extract(xs[r]).astype(np.complex128)
            e_b = _rec_err(rm, rb)
            xg, itg = ses.solve(bs[r], tols[r], niter=6000, algopt=2, nrestart=10)
            if itg < 6000:
                ngmres += 1
                res = np.linalg.norm(bs[r] - ses.fdfd_apply(xg)) / np.linalg.norm(bs[r])
                assert res < 2 * tols[r], (f, res)
                rg = ses.extract(xg).astype(np.complex128)
                e_g, e_bg = _rec_err(rm, rg), _rec_err(rg, rb)
            else:
                e_g = e_bg = 0.0
            print("f=%g iterations mixed %d bicgstab %d gmres %d; receiver error vs bicgstab %.2e, vs "
                  "gmres %.2e, gmres vs bicgstab %.2e" % (f, its[r], itb, itg, e_b, e_g, e_bg), flush=True)
            assert max(e_b, e_g, e_bg) < 1e-6, (f, e_b, e_g, e_bg)
        assert ngmres >= 2
        # cut short after one refinement step, finished by the serving
        # engine's FP64 continuation
        f, b, tol = freqs[0], bs[0], tols[0]
        (x1,), _, (rel1,), n1 = ses.solve_mixed([b], [f], [tol], _lib.MixedOpts.default(max_outer=1))
        assert n1 == 0 and tol < rel1 <= 1.0 and np.all(np.isfinite(x1)), rel1
        L = _lib.lib()
        dx, db = C.c_void_p(), C.c_void_p()
        for d, a in ((dx, x1), (db, b)):
            a = np.ascontiguousarray(a, np.complex128)
            _lib.check(L.csem_malloc(C.byref(d), a.nbytes), "malloc")
            _lib.check(L.csem_memcpy_h2d(d, a.ctypes.data, a.nbytes, None), "h2d")
        try:
            itc, relc = serve.fp64_continue(ses.model, f, dx, db, tol, rel1, 6000, threading.Lock())
            xc = np.zeros(ses.n, np.complex128)
            _lib.check(L.csem_memcpy_d2h(xc.ctypes.data, dx, xc.nbytes, None), "d2h")
        finally:
            L.csem_free(dx)
            L.csem_free(db)
        assert 0 <= itc < 6000 and relc < tol, (f, itc, relc)
        ses.set_freq(f)
        res = np.linalg.norm(b - ses.fdfd_apply(xc)) / np.linalg.norm(b)
        assert res < 2 * tol, (f, res)
        e = _rec_err(ses.extract(xc).astype(np.complex128), rbs[0])
        print("f=%g mixed cut at relres %.2e; FP64 continued %d iterations, receiver error %.2e"
              % (f, rel1, itc, e), flush=True)
        assert e < 1e-6, (f, e)
    finally:
        ses.close()


@pytest.mark.gpu
def test_vti_refinement_guard_without_the_stall_cut(workdir):
    """With the stagnation cut off (stall_iters < 0) the FP32 inner solve
    at 0.25 Hz on this model runs its whole budget and a refinement step
    fails to lower the true residual: the guard must undo it and stop that
    right-hand side with its best solution -- finite, relres below 1 and
    equal to the true residual of the x returned -- instead of diverging."""
    from demo_matrixfree_csem_amd import _lib
    from demo_matrixfree_csem_amd.session import Session
    s = surveys.vti_marine(2, 36, 24, 24, 33)
    surveys.write_files(workdir, s)
    f = 0.25
    print("case vti_marine 36x24x24 refinement guard, stall cut off", flush=True)
    ses = Session(s.argv(workdir))
    try:
        ses.set_freq(f)
        b, tol = ses.inject(), _tol(f)
        (x,), (it,), (rel,), n = ses.solve_mixed([b], [f], [tol],
                                                 _lib.MixedOpts.default(max_inner=6000, stall_iters=-1))
        print("mixed without the cut: %d inner iterations, relres %.2e" % (it, rel), flush=True)
        assert n == 0 and tol < rel < 1e-3 and np.all(np.isfinite(x)), (n, it, rel)
        res = np.linalg.norm(b - ses.fdfd_apply(x)) / np.linalg.norm(b)
        assert abs(res - rel) <= 1e-3 * rel, (res, rel)
    finally:
        ses.close()


@pytest.mark.gpu
@pytest.mark.parametrize("stall", [0, 16], ids=["default-cut", "forced-cuts"])
def test_vti_second_model_stall_handling(workdir, stall):
    """The stagnation handling on an anisotropic model the threshold was
    not tuned on (seed 5): with the default cut, and with a cut after 16
    iterations without progress (which makes the refinement go through its
    candidates -- the cut solve's last iterate, its best iterate, then
    FGMRES cycles), every right-hand side either reaches tol with receivers
    within 1e-6 of FP64 F-BiCGStab's, or stops with a finite solution of
    relative residual at most 1 that the FP64 continuation finishes to the
    same receivers."""
    import ctypes as C
    import threading
    from demo_matrixfree_csem_amd import _lib, serve
    from demo_matrixfree_csem_amd.session import Session
    s = surveys.vti_marine(5, 36, 24, 24, 33)
    surveys.write_files(workdir, s)
    freqs = [0.25, 0.75]
    print("case vti_marine seed 5, stall_iters %d" % stall, flush=True)
    ses = Session(s.argv(workdir))
    L = _lib.lib()
    try:
        bs, tols = [], []
        for f in freqs:
            ses.set_freq(f)
            bs.append(ses.inject())
            tols.append(_tol(f))
        xs, its, rel, nconv = ses.solve_mixed(bs, freqs, tols, _lib.MixedOpts.default(max_inner=4000,
                                                                                     stall_iters=stall))
        print("mixed: iterations %s, relres %s, converged %d" % (list(its), list(rel), nconv), flush=True)
        for r, f in enumerate(freqs):
            assert np.all(np.isfinite(xs[r])) and rel[r] <= 1.0
            x = xs[r]
            if not rel[r] < tols[r]:
                dx, db = C.c_void_p(), C.c_void_p()
                for d, a in ((dx, x), (db, bs[r])):
                    a = np.ascontiguousarray(a, np.complex128)
                    _lib.check(L.csem_malloc(C.byref(d), a.nbytes), "malloc")
                    _lib.check(L.csem_memcpy_h2d(d, a.ctypes.data, a.nbytes, None), "h2d")
                try:
                    itc, relc = serve.fp64_continue(ses.model, f, dx, db, tols[r], rel[r], 6000, threading.Lock())
                    x = np.zeros(ses.n, np.complex128)
                    _lib.check(L.csem_memcpy_d2h(x.ctypes.data, dx, x.nbytes, None), "d2h")
                finally:
                    L.csem_free(dx)
                    L.csem_free(db)
                print("f=%g continued in FP64: %d iterations, relres %.2e" % (f, itc, relc), flush=True)
                assert relc < tols[r]
            ses.set_freq(f)
            res = np.linalg.norm(bs[r] - ses.fdfd_apply(x)) / np.linalg.norm(bs[r])
            assert res < 2 * tols[r], (f, res)
            xb, itb = ses.solve(bs[r], tols[r], niter=6000, algopt=1)
            assert itb < 6000
            e = _rec_err(ses.extract(x).astype(np.complex128), ses.extract(xb).astype(np.complex128))
            print("f=%g: FP64 F-BiCGStab %d iterations; receiver error %.2e" % (f, itb, e), flush=True)
            assert e < 1e-6, (f, e)
    finally:
        ses.close()
