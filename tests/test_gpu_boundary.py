"""The drop-in boundary on the device, through the C-ABI.

* aphi_apply (fdfd.c:513-821, preco=3) and gmres_apply (gmres_init_free.c,
  preco=2) equal the oracle bit for bit; the oracle equals the reference
  build (tests/test_oracle_ref.py).
* F-BiCGStab / FGMRES with those preconditioners: in the reference's
  serial summation order the device solution equals the reference build's
  own (golden ref_tiny.npz) bit for bit; in the device's chunked order it
  equals the oracle's in the same order.
* The Krylov drivers take any op_t: a caller's host function is called on
  host copies, and gives the same solution as the device operator it
  wraps; a solve over host operators alone needs no model.
* The reference's own main.c, unchanged and linked against libcsem.so
  (oracle/_ref/ref_main_csem, built where the reference sources are),
  writes the same emf_0001.txt as this package's csem_main, and both equal
  the oracle's receivers.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from demo_matrixfree_csem_amd import surveys, _lib
from demo_matrixfree_csem_amd.session import Session
import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _rand(n, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal(n) + 1j * r.standard_normal(n)


def _pair(workdir, s, nfft=64):
    s.flags.update(n1fft=nfft, n2fft=nfft)
    surveys.write_files(workdir, s)
    ses = Session(s.argv(workdir))
    m = s.model
    orc = O.Oracle(m.x1node, m.x2node, m.x3node, m.rho11, m.rho22, m.rho33, nb=s.flags["nb"],
                   airbc=s.flags["airbc"], fvm=s.flags.get("fvm", 0), nfft=nfft,
                   rho_air=float(s.flags.get("rho_air", 1e6)))
    return ses, orc


@pytest.mark.parametrize("airbc,fvm,big", [(0, 0, False), (1, 0, False), (1, 1, False), (0, 1, False), (1, 0, True)])
def test_aphi_bitwise(workdir, airbc, fvm, big):
    print("case aphi airbc=%d fvm=%d big=%d" % (airbc, fvm, big), flush=True)
    kw = dict(nx=40, ny=33, nz=25) if big else {}
    s = surveys.tiny(8, airbc=airbc, **kw)
    s.flags["fvm"] = fvm
    ses, orc = _pair(workdir, s)
    try:
        for f in (0.25, 1.25):
            ses.set_freq(f)
            orc.set_freq(f)
            x = _rand(ses.n, 3)
            y = ses.aphi_apply(x)
            yo = orc.aphi(x)
            assert np.array_equal(y, yo), np.abs(y - yo).max() / np.abs(yo).max()
    finally:
        ses.close()


def test_gmres_apply_bitwise(workdir):
    print("case gmres_apply mp=8", flush=True)
    s = surveys.tiny(9, airbc=1)
    ses, orc = _pair(workdir, s)
    try:
        ses.set_freq(0.75)
        orc.set_freq(0.75)
        orc.set_red(O.RED_CHUNKED)
        x = _rand(ses.n, 5)
        y = ses.gmres_apply(x, 8)
        yo = orc.gmres_prec(x, mp=8)
        assert np.array_equal(y, yo), np.abs(y - yo).max() / np.abs(yo).max()
    finally:
        ses.close()


def _golden_case():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg.CASES[0]


@pytest.mark.parametrize("preco,algopt", [(2, 1), (2, 2), (3, 1), (3, 2)])
def test_other_preconditioner_solves(workdir, preco, algopt):
    """60 iterations from x = 0 at 0.75 Hz with mp = 8 (main.c:104-118)."""
    print("case solve preco=%d algopt=%d" % (preco, algopt), flush=True)
    g = np.load(os.path.join(GOLDEN, "ref_tiny.npz"))
    name, kw, fl = _golden_case()
    s = surveys.tiny(**kw)
    s.flags.update(fl)
    ses, orc = _pair(workdir, s)
    L = _lib.lib()
    try:
        f = 0.75
        ses.set_freq(f)
        orc.set_freq(f)
        orc.set_mp(8)
        tol = float(np.float32(np.float32(1e-8) / np.float32(f)))
        b = ses.inject()
        L.csem_set_reduction(1)
        try:
            x, it = ses.solve(b, tol, niter=60, algopt=algopt, preco=preco, nrestart=10, mp=8)
        finally:
            L.csem_set_reduction(0)
        ref = g["%s_solve_p%d_a%d" % (name, preco, algopt)]
        assert np.array_equal(x, ref), np.abs(x - ref).max() / np.abs(ref).max()
        x, it = ses.solve(b, tol, niter=60, algopt=algopt, preco=preco, nrestart=10, mp=8)
        orc.set_red(O.RED_CHUNKED)
        if algopt == 1:
            xo, ito, _ = orc.fbicgstab(b, 60, tol, precond=preco)
        else:
            xo, ito, _ = orc.fgmres(b, 6, tol, 10, precond=preco)
        assert it == ito and np.array_equal(x, xo)
    finally:
        ses.close()


OP_T = C.CFUNCTYPE(None, C.c_int, C.c_void_p, C.c_void_p)


def test_host_operator_through_copies(workdir):
    """fbicgstab with a caller's op_t (here a host function that itself calls
    fdfd_apply on host arrays) equals fbicgstab with fdfd_apply, bit for bit;
    gmres_init on a caller's op_t likewise."""
    print("case host op_t", flush=True)
    s = surveys.tiny(10, airbc=1)
    ses, _ = _pair(workdir, s)
    L = _lib.lib()
    calls = []

    def my_apply(n, x, y):
        calls.append(n)
        L.fdfd_apply(n, x, y)

    cb = OP_T(my_apply)
    try:
        ses.set_freq(1.25)
        b = ses.inject()
        tol = float(np.float32(np.float32(1e-8) / np.float32(1.25)))
        x0, it0 = ses.solve(b, tol, niter=80, algopt=1, preco=1)
        n = ses.n
        x = np.zeros(n, np.complex128)
        L.fbicgstab(n, x.ctypes.data, b.ctypes.data, C.cast(cb, C.c_void_p), _lib.fn_addr("lusgs_apply"), 80, tol, 0)
        assert L.csem_last_status() == 0
        assert L.csem_last_iterations() == it0 and np.array_equal(x, x0)
        assert calls and set(calls) == {n}
        # the inner-GMRES preconditioner on the caller's operator
        y0 = ses.gmres_apply(b, 4)
        L.gmres_init(n, 4, C.cast(cb, C.c_void_p))
        y = np.zeros(n, np.complex128)
        L.gmres_apply(n, b.ctypes.data, y.ctypes.data)
        L.gmres_free()
        assert np.array_equal(y, y0)
    finally:
        ses.close()


def test_solve_over_host_operators_only():
    """No model is needed when every operator is the caller's: BiCGStab on
    a small dense complex system converges to numpy's solution."""
    print("case host-only solve", flush=True)
    r = np.random.default_rng(4)
    n = 60
    A = np.diag(4 + r.uniform(0, 1, n) + 1j * r.uniform(0, 1, n)) + 0.3 * (r.standard_normal((n, n)) +
                                                                         1j * r.standard_normal((n, n))) / np.sqrt(n)
    L = _lib.lib()

    def apply(nn, xp, yp):
        x = np.ctypeslib.as_array(C.cast(xp, C.POINTER(C.c_double)), (2 * nn,)).view(np.complex128)
        y = np.ctypeslib.as_array(C.cast(yp, C.POINTER(C.c_double)), (2 * nn,)).view(np.complex128)
        y[:] = A @ x

    cb = OP_T(apply)
    b = r.standard_normal(n) + 1j * r.standard_normal(n)
    x = np.zeros(n, np.complex128)
    L.bicgstab(n, x.ctypes.data, b.ctypes.data, C.cast(cb, C.c_void_p), 200, 1e-12, 0)
    assert L.csem_last_status() == 0
    assert np.allclose(x, np.linalg.solve(A, b), rtol=0, atol=1e-9 * np.abs(b).max())


def _run(cmd, cwd):
    print("run", " ".join(cmd[:1]), flush=True)
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    return r.stdout


def _emf_rows(path):
    with open(path) as f:
        return [l for l in f.read().splitlines()[1:] if l.strip()]


REF_MAIN = os.path.join(ROOT, "oracle", "_ref", "ref_main_csem")
CSEM_MAIN = os.path.join(ROOT, "demo_matrixfree_csem_amd", "csem_main")


@pytest.mark.skipif(not os.path.exists(REF_MAIN), reason="reference main.c linked against libcsem.so not built")
@pytest.mark.parametrize("preco,algopt", [(1, 1), (1, 2), (2, 1), (3, 1)])
def test_reference_main_drives_the_library(workdir, preco, algopt):
    print("case reference main.c preco=%d algopt=%d" % (preco, algopt), flush=True)
    s = surveys.tiny(11, airbc=1)
    s.flags.update(n1fft=64, n2fft=64, preco=preco, algopt=algopt, niter=300, mp=6)
    s.freqs = [0.5, 1.25]
    surveys.write_files(workdir, s)
    argv = s.argv(workdir)
    d1, d2 = os.path.join(workdir, "ref"), os.path.join(workdir, "own")
    os.makedirs(d1)
    os.makedirs(d2)
    _run([REF_MAIN] + argv, d1)
    _run([CSEM_MAIN] + argv, d2)
    rows = _emf_rows(os.path.join(d1, "emf_0001.txt"))
    assert rows == _emf_rows(os.path.join(d2, "emf_0001.txt"))
    # the oracle's receivers (device summation order), float32 as written
    m = s.model
    orc = O.Oracle(m.x1node, m.x2node, m.x3node, m.rho11, m.rho22, m.rho33, nb=s.flags["nb"], airbc=1, nfft=64)
    orc.set_red(O.RED_CHUNKED)
    orc.set_mp(6)
    vals = {}
    for ifr, f in enumerate(s.freqs):
        orc.set_freq(f)
        tol = float(np.float32(np.float32(1e-8) / np.float32(f)))
        b = orc.inject(0, *s.acq.sources[0][:3])
        if algopt == 1:
            x, _, _ = orc.fbicgstab(b, 300, tol, precond=preco)
        else:
            x, _, _ = orc.fgmres(b, 30, tol, 10, precond=preco)
        for ch in range(3):
            for ir, rr in enumerate(s.acq.receivers):
                vals[(ir + 1, ch + 1, ifr + 1)] = np.complex64(orc.extract(ch, *rr[:3], x))
    assert len(rows) == len(vals)
    for l in rows:
        t = l.split()
        v = vals[(int(t[1]), int(t[2]), int(t[3]))]
        assert (float(t[4]), float(t[5])) == (float("%e" % v.real), float("%e" % v.imag)), l
