"""Batched mixed-precision path on the GPU, checked against the FP64
oracle (the pinned CPU restatement of the reference) and against the FP64
device kernels.

Tolerances (stated here and in the assertions):
  * FP32 operator vs FP64 operator: norm-wise relative 2e-6 on random
    fields (FP32 rounding of inputs and of the curl-curl differences);
  * FP32 LU-SGS vs FP64 LU-SGS: norm-wise relative 1e-4 (a preconditioner:
    only its action needs to be close);
  * mixed solve (with either FP32 preconditioner: the lexicographic LU-SGS
    or its red-black ordering) vs the oracle's FP64 F-BiCGStab, both to
    tol = 1e-8/f:
    every receiver value within 1e-6 of the channel's largest value, and
    the true FP64 residual |b - A x| / |b| (oracle operator) below tol;
  * batched vs alone: the same receivers within 1e-6 (channel-normalised);
  * graph replay vs eager launches: bit-identical.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import oracle as O  # noqa: E402

from demo_matrixfree_csem_amd import surveys, _lib  # noqa: E402
from demo_matrixfree_csem_amd.session import Session, freq_consts  # noqa: E402

pytestmark = pytest.mark.gpu
FREQS = [0.25, 0.75, 1.25]


def _setup(workdir, seed=2, **kw):
    s = surveys.tiny(seed, airbc=1, **kw)
    s.flags.update(n1fft=64, n2fft=64, chrec="Ex,Ey,Ez,Hx,Hy,Hz")
    s.freqs = list(FREQS)
    surveys.write_files(workdir, s)
    ses = Session(s.argv(workdir))
    m = s.model
    orc = O.Oracle(m.x1node, m.x2node, m.x3node, m.rho11, m.rho22, m.rho33, nb=s.flags["nb"], airbc=1, nfft=64)
    return s, ses, orc


def _dev(a):
    d = C.c_void_p()
    _lib.check(_lib.lib().csem_malloc(C.byref(d), a.nbytes), "malloc")
    _lib.check(_lib.lib().csem_memcpy_h2d(d, a.ctypes.data, a.nbytes, None), "h2d")
    return d


def _host(d, like):
    out = np.empty_like(like)
    _lib.check(_lib.lib().csem_memcpy_d2h(out.ctypes.data, d, out.nbytes, None), "d2h")
    _lib.check(_lib.lib().csem_sync(None), "sync")
    return out


def _tol(f):
    return float(np.float32(np.float32(1e-8) / np.float32(f)))


def test_apply32_close_to_fp64(workdir):
    print("case apply32", flush=True)
    s, ses, orc = _setup(workdir)
    L = _lib.lib()
    try:
        n, R = ses.n, 3
        rng = np.random.default_rng(11)
        x = (rng.standard_normal((R, n)) + 1j * rng.standard_normal((R, n))).astype(np.complex64)
        fr = (_lib.FreqT * R)(*[freq_consts(f) for f in FREQS])
        dx, dy = _dev(x), _dev(np.zeros_like(x))
        _lib.check(L.csem_apply32(ses.model, R, fr, dx, dy, None), "apply32")
        y32 = _host(dy, x)
        for r, f in enumerate(FREQS):
            orc.set_freq(f)
            y64 = orc.apply(x[r].astype(np.complex128))
            err = np.linalg.norm(y32[r] - y64) / np.linalg.norm(y64)
            assert err < 2e-6, (f, err)
        # the same right-hand side alone gives the same values
        dx1, dy1 = _dev(x[1:2].copy()), _dev(np.zeros_like(x[1:2]))
        _lib.check(L.csem_apply32(ses.model, 1, (_lib.FreqT * 1)(fr[1]), dx1, dy1, None), "apply32")
        y1 = _host(dy1, x[1:2])
        assert np.abs(y1[0] - y32[1]).max() <= 1e-6 * np.abs(y32[1]).max()
        for d in (dx, dy, dx1, dy1):
            L.csem_free(d)
    finally:
        ses.close()


REPEATABLE = ["k_apply", "k_hz0", "k_air_rows_fwd", "k_air_cols", "k_air_ghost", "k_lu_pre", "k_lu_smith",
              "k_sgs_ex<fwd>", "k_sgs_ey<fwd>", "k_sgs_ez<fwd>", "k_sgs_ez<bwd>", "k_sgs_ey<bwd>", "k_sgs_ex<bwd>",
              "k_mx_apply", "k_mx_hz0", "k_mx_air_rows", "k_mx_air_cols", "k_mx_air_ghost", "k_mx_lu_pre",
              "k_mx_sgs", "k_mx_unskew"]


def test_repeat_loops_leave_results_unchanged(workdir):
    """The timing loops of csem_kernel_repeat launch a kernel again with the
    same arguments: with loops on every repeatable kernel (and every launch
    timed through hipExtLaunchKernelGGL) the FP64 operator and LU-SGS still
    equal the oracle bit for bit and the FP32 ones equal a run without
    timers; every loop's launches are counted; a kernel whose output feeds
    its input is refused."""
    print("case repeat loops", flush=True)
    s, ses, orc = _setup(workdir)
    L = _lib.lib()
    try:
        n, R = ses.n, 3
        rng = np.random.default_rng(13)
        x = (rng.standard_normal((R, n)) + 1j * rng.standard_normal((R, n))).astype(np.complex64)
        fr = (_lib.FreqT * R)(*[freq_consts(f) for f in FREQS])
        dx, dy, dz = _dev(x), _dev(np.zeros_like(x)), _dev(np.zeros_like(x))
        _lib.check(L.csem_apply32(ses.model, R, fr, dx, dy, None), "apply32")
        _lib.check(L.csem_lusgs32(ses.model, R, fr, dx, dz, None), "lusgs32")
        y0, z0 = _host(dy, x), _host(dz, x)
        L.csem_kernel_timing(1)
        try:
            for k in REPEATABLE:
                _lib.check(L.csem_kernel_repeat(_lib.kernel_id(k), 3, 2), "repeat " + k)
            assert L.csem_kernel_repeat(_lib.kernel_id("k_mx_krylov"), 3, 2) == -1
            assert L.csem_kernel_repeat(_lib.kernel_id("k_air_rows_inv"), 3, 2) == -1
            ses.set_freq(0.75)
            orc.set_freq(0.75)
            x64 = x[1].astype(np.complex128)
            assert np.array_equal(ses.fdfd_apply(x64), orc.apply(x64))
            assert np.array_equal(ses.lusgs_apply(x64), orc.lusgs(x64))
            _lib.check(L.csem_apply32(ses.model, R, fr, dx, dy, None), "apply32")
            _lib.check(L.csem_lusgs32(ses.model, R, fr, dx, dz, None), "lusgs32")
            y1, z1 = _host(dy, x), _host(dz, x)
            for k in REPEATABLE:
                ms, cnt = _lib.kernel_loop_stats(k)
                assert cnt > 0 and cnt % 3 == 0 and ms > 0, (k, cnt, ms)
        finally:
            L.csem_kernel_timing(0)
        assert np.array_equal(y1, y0) and np.array_equal(z1, z0)
        for d in (dx, dy, dz):
            L.csem_free(d)
    finally:
        ses.close()


def test_apply32_every_tile(workdir):
    """Every tile shape the autotuner may pick gives the same operator."""
    print("case apply32 tiles", flush=True)
    s, ses, orc = _setup(workdir)
    L = _lib.lib()
    try:
        n = ses.n
        rng = np.random.default_rng(13)
        for R in (1, 3, 4):
            x = (rng.standard_normal((R, n)) + 1j * rng.standard_normal((R, n))).astype(np.complex64)
            fq = (FREQS * 2)[:R]
            fr = (_lib.FreqT * R)(*[freq_consts(f) for f in fq])
            dx, dy = _dev(x), _dev(np.zeros_like(x))
            ref = []
            for r, f in enumerate(fq):
                orc.set_freq(f)
                ref.append(orc.apply(x[r].astype(np.complex128)))
            first = None
            for rows in ((4, 6) if R <= 3 else (2, 4)):
                for planes in (8, 10, 12, 16):
                    _lib.check(L.csem_mixed_set_tile(ses.model, R, rows, planes), "set_tile")
                    _lib.check(L.csem_apply32(ses.model, R, fr, dx, dy, None), "apply32")
                    y = _host(dy, x)
                    for r in range(R):
                        err = np.linalg.norm(y[r] - ref[r]) / np.linalg.norm(ref[r])
                        assert err < 2e-6, (R, rows, planes, r, err)
                    # the tile changes no arithmetic: every tile gives the same bits
                    if first is None:
                        first = y
                    assert np.array_equal(y, first), (R, rows, planes)
            L.csem_free(dx)
            L.csem_free(dy)
    finally:
        ses.close()


def test_lusgs32_close_to_fp64(workdir):
    print("case lusgs32", flush=True)
    s, ses, orc = _setup(workdir)
    L = _lib.lib()
    try:
        n, R = ses.n, 3
        rng = np.random.default_rng(12)
        x = (rng.standard_normal((R, n)) + 1j * rng.standard_normal((R, n))).astype(np.complex64)
        fr = (_lib.FreqT * R)(*[freq_consts(f) for f in FREQS])
        dx, dy = _dev(x), _dev(np.zeros_like(x))
        _lib.check(L.csem_lusgs32(ses.model, R, fr, dx, dy, None), "lusgs32")
        y32 = _host(dy, x)
        for r, f in enumerate(FREQS):
            orc.set_freq(f)
            y64 = orc.lusgs(x[r].astype(np.complex128))
            err = np.linalg.norm(y32[r] - y64) / np.linalg.norm(y64)
            assert err < 1e-4, (f, err)
        L.csem_free(dx)
        L.csem_free(dy)
    finally:
        ses.close()


def test_lusgs32_follows_frequency_changes(workdir):
    """The FP32 records' per-frequency half (h = iwmu/aii) is kept per
    solve slot and rebuilt when a right-hand side's frequency changes: the
    preconditioner at 0.25, 1.25 and again 0.25 Hz on one slot gives the
    first result bit for bit the third time and matches the oracle at each
    frequency (an h left from the other frequency misses by far more)."""
    print("case lusgs32 frequency changes", flush=True)
    s, ses, orc = _setup(workdir)
    L = _lib.lib()
    try:
        n = ses.n
        rng = np.random.default_rng(21)
        x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
        dx, dy = _dev(x), _dev(np.zeros_like(x))
        ys = []
        for f in (0.25, 1.25, 0.25):
            fr = (_lib.FreqT * 1)(freq_consts(f))
            _lib.check(L.csem_lusgs32(ses.model, 1, fr, dx, dy, None), "lusgs32")
            ys.append(_host(dy, x))
            orc.set_freq(f)
            y64 = orc.lusgs(x.astype(np.complex128))
            err = np.linalg.norm(ys[-1] - y64) / np.linalg.norm(y64)
            print("f=%g lusgs32 vs oracle %.2e" % (f, err), flush=True)
            assert err < 1e-4, (f, err)
        assert np.array_equal(ys[0].view(np.float32), ys[2].view(np.float32))
        assert not np.array_equal(ys[0].view(np.float32), ys[1].view(np.float32))
        L.csem_free(dx)
        L.csem_free(dy)
    finally:
        ses.close()


def test_fp32_kernels_multi_tile(workdir):
    """Several operator tiles along x and y, LU-SGS slices of more than 64
    rows (several waves per slice, lagged edge exchange): the FP32 operator
    and LU-SGS stay within their tolerances of the FP64 oracle."""
    print("case fp32 multi-tile", flush=True)
    s = surveys.tiny(5, airbc=1, nx=130, ny=75, nz=70)
    s.flags.update(n1fft=128, n2fft=128)
    s.freqs = list(FREQS)
    surveys.write_files(workdir, s)
    ses = Session(s.argv(workdir))
    m = s.model
    orc = O.Oracle(m.x1node, m.x2node, m.x3node, m.rho11, m.rho22, m.rho33, nb=s.flags["nb"], airbc=1, nfft=128)
    L = _lib.lib()
    try:
        n, R = ses.n, 3
        rng = np.random.default_rng(17)
        x = (rng.standard_normal((R, n)) + 1j * rng.standard_normal((R, n))).astype(np.complex64)
        fr = (_lib.FreqT * R)(*[freq_consts(f) for f in FREQS])
        dx, dy = _dev(x), _dev(np.zeros_like(x))
        _lib.check(L.csem_apply32(ses.model, R, fr, dx, dy, None), "apply32")
        ya = _host(dy, x)
        _lib.check(L.csem_lusgs32(ses.model, R, fr, dx, dy, None), "lusgs32")
        yl = _host(dy, x)
        for r, f in enumerate(FREQS):
            orc.set_freq(f)
            xr = x[r].astype(np.complex128)
            ya64, yl64 = orc.apply(xr), orc.lusgs(xr)
            ea = np.linalg.norm(ya[r] - ya64) / np.linalg.norm(ya64)
            el = np.linalg.norm(yl[r] - yl64) / np.linalg.norm(yl64)
            print("f=%g apply32 %.2e lusgs32 %.2e" % (f, ea, el), flush=True)
            assert ea < 2e-6, (f, ea)
            assert el < 1e-4, (f, el)
        L.csem_free(dx)
        L.csem_free(dy)
    finally:
        ses.close()


@pytest.mark.parametrize("shape", [dict(), dict(nx=130, ny=75, nz=70)], ids=["tiny", "multi-wave"])
def test_fp32_kernels_same_alone_and_batched(workdir, shape):
    """Each right-hand side's FP32 operator and LU-SGS result is the same,
    bit for bit, whether it is applied alone (R=1) or in a batch of 2, 3 or
    4: no kernel's arithmetic depends on the batch size (or on the operator
    tile the autotuner picks for it), so a solve's path does not either."""
    print("case fp32 alone vs batched %s" % shape, flush=True)
    s = surveys.tiny(4, airbc=1, **shape)
    s.flags.update(n1fft=64, n2fft=64)
    s.freqs = [0.375, 0.75, 1.25, 0.25]
    surveys.write_files(workdir, s)
    ses = Session(s.argv(workdir))
    L = _lib.lib()
    try:
        n = ses.n
        rng = np.random.default_rng(21)
        x = (rng.standard_normal((4, n)) + 1j * rng.standard_normal((4, n))).astype(np.complex64)
        for name, fn in (("apply32", L.csem_apply32), ("lusgs32", L.csem_lusgs32)):
            alone = []
            for r in range(4):
                fr = (_lib.FreqT * 1)(freq_consts(s.freqs[r]))
                dx, dy = _dev(x[r:r + 1].copy()), _dev(np.zeros_like(x[r:r + 1]))
                _lib.check(fn(ses.model, 1, fr, dx, dy, None), name)
                alone.append(_host(dy, x[r:r + 1])[0])
                L.csem_free(dx)
                L.csem_free(dy)
            for R in (2, 3, 4):
                fr = (_lib.FreqT * R)(*[freq_consts(f) for f in s.freqs[:R]])
                dx, dy = _dev(x[:R].copy()), _dev(np.zeros_like(x[:R]))
                _lib.check(fn(ses.model, R, fr, dx, dy, None), name)
                y = _host(dy, x[:R])
                L.csem_free(dx)
                L.csem_free(dy)
                for r in range(R):
                    assert np.array_equal(y[r], alone[r]), (name, R, r, float(np.abs(y[r] - alone[r]).max()))
    finally:
        ses.close()


@pytest.mark.parametrize("shape", [dict(), dict(nx=130, ny=75, nz=70)], ids=["tiny", "multi-wave"])
def test_lu_tiles_and_fft_columns_change_nothing(workdir, shape):
    """Every LU-SGS pre-pass tile and every FFT column-pass width the
    autotuner may pick gives the same FP32 LU-SGS and operator, bit for bit,
    at one and three right-hand sides; a solve leaves its picks readable."""
    print("case lu tiles / fft columns %s" % shape, flush=True)
    s = surveys.tiny(4, airbc=1, **shape)
    s.flags.update(n1fft=64, n2fft=64)
    s.freqs = [0.375, 0.75, 1.25]
    surveys.write_files(workdir, s)
    ses = Session(s.argv(workdir))
    L = _lib.lib()
    try:
        n = ses.n
        rng = np.random.default_rng(23)
        for R in (1, 3):
            x = (rng.standard_normal((R, n)) + 1j * rng.standard_normal((R, n))).astype(np.complex64)
            fr = (_lib.FreqT * R)(*[freq_consts(f) for f in s.freqs[:R]])
            dx, dy = _dev(x), _dev(np.zeros_like(x))
            base = {}
            try:
                for v in range(3):
                    for cols in (16, 32, 64):  # nfft 64: 64 * cols in [1024, 4096]
                        pre = (C.c_int * 6)(*([v] * 6))
                        _lib.check(L.csem_mixed_set_lu_tiles(ses.model, R, pre, cols), "set_lu_tiles")
                        for name, fn in (("lusgs32", L.csem_lusgs32), ("apply32", L.csem_apply32)):
                            _lib.check(fn(ses.model, R, fr, dx, dy, None), name)
                            y = _host(dy, x)
                            if name not in base:
                                base[name] = y
                            assert np.array_equal(y, base[name]), (R, v, cols, name)
                bad = (C.c_int * 6)(*([3] * 6))
                assert L.csem_mixed_set_lu_tiles(ses.model, R, bad, 32) < 0
                assert L.csem_mixed_set_lu_tiles(ses.model, R, pre, 8) < 0
            finally:
                L.csem_free(dx)
                L.csem_free(dy)
        # a solve at a batch size not fixed above times the candidates
        ses.set_freq(s.freqs[0])
        b = ses.inject()
        ses.solve_mixed([b, b], s.freqs[:2], [1e-3, 1e-3])
        pre, cols = (C.c_int * 6)(), C.c_int()
        assert L.csem_mixed_tuning_lu(ses.model, 2, pre, C.byref(cols)) == 0
        print("tuned pre-pass tiles %s, FFT columns per block %d" % (list(pre), cols.value), flush=True)
        assert all(0 <= v < 3 for v in pre) and cols.value in (16, 32, 64)
    finally:
        ses.close()


def _receivers(ses, x, f):
    ses.set_freq(f)  # the H channels scale by 1 / (i w mu0)
    return ses.extract(x).astype(np.complex128)


def _rec_err(a, b):
    """max over channels of max |a - b| / max |b| (channel-normalised)."""
    scale = np.abs(b).max(axis=1, keepdims=True)
    return float((np.abs(a - b) / np.where(scale > 0, scale, 1)).max())


@pytest.mark.parametrize("precond", [_lib.PRECOND_LUSGS, _lib.PRECOND_RBSGS])
def test_mixed_solve_matches_fp64_oracle(workdir, precond):
    print("case mixed vs oracle, preconditioner %d" % precond, flush=True)
    s, ses, orc = _setup(workdir)
    try:
        bs, tols = [], []
        for f in FREQS:
            ses.set_freq(f)
            bs.append(ses.inject())
            tols.append(_tol(f))
        xs, its, rel, nconv = ses.solve_mixed(bs, FREQS, tols, _lib.MixedOpts.default(precond=precond, max_inner=4000))
        assert nconv == 3, (its, rel)
        for r, f in enumerate(FREQS):
            orc.set_freq(f)
            orc.set_red(O.RED_CHUNKED)
            xo, ito, _ = orc.fbicgstab(bs[r], 2000, tols[r])
            res = np.linalg.norm(bs[r] - orc.apply(xs[r])) / np.linalg.norm(bs[r])
            assert res < tols[r], (f, res)
            e = _rec_err(_receivers(ses, xs[r], f), _receivers(ses, xo, f))
            print("f=%g inner iterations %d (fp64 %d) true relres %.2e receiver error %.2e"
                  % (f, its[r], ito, res, e), flush=True)
            assert e < 1e-6, (f, e)
    finally:
        ses.close()


def test_mixed_batch_equals_alone(workdir):
    print("case mixed batch vs alone", flush=True)
    s, ses, orc = _setup(workdir, seed=3)
    try:
        bs, tols = [], []
        for f in FREQS:
            ses.set_freq(f)
            bs.append(ses.inject())
            tols.append(_tol(f))
        xs, its, rel, nconv = ses.solve_mixed(bs, FREQS, tols)
        assert nconv == 3
        for r, f in enumerate(FREQS):
            x1, it1, rel1, nc1 = ses.solve_mixed([bs[r]], [f], [tols[r]])
            assert nc1 == 1
            e = _rec_err(_receivers(ses, xs[r], f), _receivers(ses, x1[0], f))
            print("f=%g batched %d alone %d iterations, receiver difference %.2e" % (f, its[r], it1[0], e),
                  flush=True)
            assert e < 1e-6, (f, e)
    finally:
        ses.close()


def test_mixed_graph_replay_equals_eager(workdir):
    print("case mixed graph vs eager", flush=True)
    s, ses, orc = _setup(workdir, seed=4)
    try:
        bs, tols = [], []
        for f in FREQS[:2]:
            ses.set_freq(f)
            bs.append(ses.inject())
            tols.append(_tol(f))
        xg, itg, relg, _ = ses.solve_mixed(bs, FREQS[:2], tols, _lib.MixedOpts.default(use_graph=1))
        xe, ite, rele, _ = ses.solve_mixed(bs, FREQS[:2], tols, _lib.MixedOpts.default(use_graph=0))
        assert list(itg) == list(ite)
        for a, b in zip(xg, xe):
            assert np.array_equal(a, b)
        # replaying again after eager work gives the same result
        xg2, itg2, _, _ = ses.solve_mixed(bs, FREQS[:2], tols, _lib.MixedOpts.default(use_graph=1))
        assert list(itg2) == list(itg)
        for a, b in zip(xg2, xg):
            assert np.array_equal(a, b)
    finally:
        ses.close()
