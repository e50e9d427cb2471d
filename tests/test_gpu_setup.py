"""Device-side model setup (csem_model_build_device: the cells, the
edge-averaged conductivities and the airwave multipliers built on the
device) against the host setup (extend_model_init + fdfd_init) on the same
input files:

* the conductivities bit for bit (extend_model.c:108-252: lateral padding,
  air on top, quirk Q2 at the bottom, the edge averages and their boundary
  copies), on the device and in the host copy that inject_source and
  extract_emf read;
* the airwave multipliers (airbc_init_free.c:69-84) within 1e-15 of their
  largest magnitude: the device's exp, sin and cos are not the host's
  libm, so the last bits may differ;
* the FP64 operator bit for bit without the airwave boundary, within
  1e-14 (relative, norm-wise) with it;
* FP64 F-BiCGStab on the same right-hand side on the two models: without
  the airwave boundary the same iterate bit for bit after 60 iterations;
  with it, converged receivers within 5 tol of the channel's largest value
  (the two operators differ by roundoff, which BiCGStab's trajectory
  amplifies up to what the tolerance allows).
"""
import ctypes as C

import numpy as np
import pytest

from demo_matrixfree_csem_amd import surveys, _lib
from demo_matrixfree_csem_amd.session import Session, freq_consts

pytestmark = pytest.mark.gpu


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


# solve: "bitwise" (60 iterations from 0, the same iterate), "converged"
# (receivers of converged solves), or None (the model converges too slowly
# at this frequency for a test: conductivities, multipliers, operator only)
CASES = [dict(seed=3, airbc=1, fvm=0, kw={}, solve="converged"), dict(seed=4, airbc=0, fvm=0, kw={}, solve="bitwise"),
         dict(seed=5, airbc=1, fvm=1, kw=dict(nx=40, ny=33, nz=25), solve=None),
         dict(seed=6, airbc=0, fvm=1, kw=dict(nx=17, ny=12, nz=9), solve="bitwise")]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "seed%d_airbc%d_fvm%d" % (c["seed"], c["airbc"], c["fvm"]))
def test_device_setup_equals_host_setup(workdir, case):
    print("case device setup %s" % case, flush=True)
    s = surveys.tiny(case["seed"], airbc=case["airbc"], **case["kw"])
    s.flags.update(fvm=case["fvm"], n1fft=64, n2fft=64)
    surveys.write_files(workdir, s)
    L = _lib.lib()
    host = Session(s.argv(workdir))
    sig_h = {w: host.sigma(w) for w in (11, 22, 33)}
    dev = Session(s.argv(workdir), device=False, setup="device")
    m_h, m_d = host.model, dev.built
    try:
        assert dev.n == host.n
        for c, w in enumerate((11, 22, 33)):
            a, b = np.empty_like(sig_h[w]), np.empty_like(sig_h[w])
            _lib.check(L.csem_model_sigma(m_h, c, a.ctypes.data), "sigma")
            _lib.check(L.csem_model_sigma(m_d, c, b.ctypes.data), "sigma")
            assert np.array_equal(a, sig_h[w]) and np.array_equal(b, sig_h[w]), w
            assert np.array_equal(dev.sigma(w), sig_h[w]), w
        if case["airbc"]:
            N = 64
            for c in (0, 1):
                a, b = np.empty(N * N, np.complex128), np.empty(N * N, np.complex128)
                _lib.check(L.csem_model_multiplier(m_h, c, a.ctypes.data), "mult")
                _lib.check(L.csem_model_multiplier(m_d, c, b.ctypes.data), "mult")
                e = np.abs(a - b).max() / np.abs(a).max()
                print("multiplier %d: max |device - host| / max |host| = %.2e" % (c, e), flush=True)
                assert e <= 1e-15, (c, e)
        # the FP64 operator on both models
        rng = np.random.default_rng(case["seed"])
        n = host.n
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        fr = freq_consts(0.75)
        dx, dy1, dy2 = _dev(x), _dev(np.zeros(n, np.complex128)), _dev(np.zeros(n, np.complex128))
        _lib.check(L.csem_apply(m_h, C.byref(fr), dx, dy1, None), "apply")
        _lib.check(L.csem_apply(m_d, C.byref(fr), dx, dy2, None), "apply")
        y1, y2 = _host(dy1, x), _host(dy2, x)
        if case["airbc"]:
            e = np.linalg.norm(y1 - y2) / np.linalg.norm(y1)
            print("operator: relative difference %.2e" % e, flush=True)
            assert e <= 1e-14, e
        else:
            assert np.array_equal(y1, y2)
        # an FP64 solve of the same right-hand side on each model
        if case["solve"] is None:
            for d in (dx, dy1, dy2):
                L.csem_free(d)
            return
        host.set_freq(0.75)
        b = host.inject()
        db = _dev(b)
        tol = float(np.float32(np.float32(1e-8) / np.float32(0.75)))
        hist = np.zeros(2002)
        sols = []
        for m, dxs in ((m_h, dy1), (m_d, dy2)):
            _lib.check(L.csem_memset(dxs, 0, n * 16, None), "memset")
            if case["solve"] == "converged":
                it = L.csem_bicgstab(m, C.byref(fr), dxs, db, 2000, tol, 1, hist.ctypes.data, None)
                assert 0 < it < 2000, it
            else:
                _lib.check(L.csem_bicgstab(m, C.byref(fr), dxs, db, 60, 0.0, 1, hist.ctypes.data, None), "bicgstab")
            sols.append(_host(dxs, b))
        if case["solve"] == "bitwise":
            assert np.array_equal(sols[0], sols[1])
        else:
            recs = [host.extract_double(v) for v in sols]
            scale = np.abs(recs[0]).max(axis=1, keepdims=True)
            e = (np.abs(recs[0] - recs[1]) / np.where(scale > 0, scale, 1)).max()
            print("FP64 solve receivers: device setup vs host setup %.2e (tol %.2e)" % (e, tol), flush=True)
            assert e <= 5 * tol, e
        for d in (dx, dy1, dy2, db):
            L.csem_free(d)
    finally:
        L.csem_model_destroy(m_d)
        host.close()
