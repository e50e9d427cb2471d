"""The FP64 device path on full-size surveys, against the reference's own
output and against the oracle's full-size solves.

* run_layered -- the reference's checked-in example (100^3, dz = 50 m,
  nb = 10, airbc = 1, 1.25 Hz, 641 receivers; real FFTW in its airwave
  transform).  The device solve must reproduce the reference-held
  emf_0001.txt (float32 printed with %e, 7 significant digits) and its
  iterate_fbicgstab.txt (166 iterations).  Our transform is not FFTW, so
  agreement is to the print resolution and the FFT roundoff, which
  BiCGStab amplifies after a few tens of iterations: the first 20 rows of
  the residual history agree to the print resolution, the iteration count
  to +-2, the Ex and Ez receivers to 1e-6 of the channel's largest value,
  and Ey -- zero by symmetry on the source's line, so solver residue --
  to 1e-6 of the largest E value.
* layered_64 (64^3, 0.25 Hz) and layered_100 (100^3, 0.25/0.75/1.25 Hz):
  the device solve in its chunked summation order equals the oracle in
  the same order bit for bit (SHA-256 of the solution vector, iteration
  count).  In the reference's serial summation order
  (csem_set_reduction(1)) the device solve of layered_64 equals the
  oracle in that order -- what the reference computes -- bit for bit.
* The two summation orders themselves: BiCGStab amplifies last-bit
  differences of its dot products, so two solves that differ only in the
  order of their sums follow different trajectories and stop at different
  points below the same relative residual tol = 1e-8/f.  Measured on the
  oracle (fixture): run_layered 166 / 166 iterations, layered_64 274 / 272,
  layered_100 at 0.75 Hz 478 / 522 and at 1.25 Hz 427 / 455 (serial /
  chunked); receivers apart by 1e-10 .. 2.5e-8 of the channel's largest
  value.  The test bounds them by what the tolerance allows: receivers
  within 5 tol of the channel's largest value (Ey, zero by symmetry on the
  line, of the largest E value), iteration counts within 10 %.

The oracle's results are fixtures (tests/golden/configs_oracle.npz,
written by tests/golden/make_configs.py): the full-size oracle solves take
minutes to an hour of CPU time.
"""
import hashlib
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from demo_matrixfree_csem_amd import surveys  # noqa: E402
from demo_matrixfree_csem_amd.session import Session  # noqa: E402
import make_configs as MC  # noqa: E402

pytestmark = pytest.mark.gpu

FIX_PATH = os.path.join(HERE, "golden", "configs_oracle.npz")
FIX = dict(np.load(FIX_PATH)) if os.path.exists(FIX_PATH) else {}


def _fix(key):
    if key not in FIX:
        pytest.skip("fixture %s not generated" % key)
    return FIX[key]


def _session(s, workdir):
    surveys.write_files(workdir, s)
    return Session(s.argv(workdir))


def _tol(s, f):
    return float(np.float32(np.float32(s.flags["tol"]) / np.float32(f)))  # main.c:94, in float


def _table(text):
    return np.array([[float(v) for v in l.split()] for l in str(text).splitlines()[1:] if l.strip()])


def test_run_layered_reproduces_reference_output(workdir):
    print("case run_layered 1.25 Hz (100^3, airbc=1, nfft=1024)", flush=True)
    g = np.load(os.path.join(HERE, "golden", "run_layered.npz"))
    s = MC.run_layered_survey()
    ses = _session(s, workdir)
    try:
        f = 1.25
        ses.set_freq(f)
        b = ses.inject()
        cwd = os.getcwd()
        os.chdir(workdir)
        try:
            x, it = ses.solve(b, _tol(s, f), niter=1000, verb=1)
            hist = _table(open("iterate_fbicgstab.txt").read())
        finally:
            os.chdir(cwd)
        rec = ses.extract(x)  # complex64 (3, 641): what write_data prints
    finally:
        ses.close()
    ref_hist = _table(g["iterate_fbicgstab_txt"])
    ref_rec = _table(g["emf_0001_txt"])
    ref = (ref_rec[:, 4] + 1j * ref_rec[:, 5]).reshape(3, -1)
    # iterations: the reference converged at iter 166 (its last history row)
    assert abs(it - int(ref_hist[-1, 0])) <= 2, (it, ref_hist[-1, 0])
    assert len(hist) == it + 1
    assert np.allclose(hist[:20, 1], ref_hist[:20, 1], rtol=2e-6, atol=0)
    emax = np.abs(ref[[0, 2]]).max()
    for ch, scale in ((0, np.abs(ref[0]).max()), (1, emax), (2, np.abs(ref[2]).max())):
        d = np.abs(rec[ch].astype(np.complex128) - ref[ch])
        print("channel %d: max |diff| / scale = %.3g" % (ch, d.max() / scale))
        assert d.max() <= 1e-6 * scale, (ch, d.max() / scale)
    # the oracle's solve in the device's order, bit for bit
    key = "run_layered_f0_chunked"
    if key + "_sha256" in FIX:
        assert it == int(FIX[key + "_iter"])
        assert hashlib.sha256(x.tobytes()).hexdigest() == str(FIX[key + "_sha256"])


def _check_case(name, s, workdir):
    ses = _session(s, workdir)
    try:
        for i, f in enumerate(s.freqs):
            print("case %s f=%g" % (name, f), flush=True)
            ses.set_freq(f)
            b = ses.inject()
            x, it = ses.solve(b, _tol(s, f), niter=int(s.flags["niter"]))
            chk = "%s_f%d_chunked" % (name, i)
            ser = "%s_f%d_serial" % (name, i)
            assert it == int(_fix(chk + "_iter")), (it, int(FIX[chk + "_iter"]))
            assert hashlib.sha256(x.tobytes()).hexdigest() == str(_fix(chk + "_sha256"))
            its = int(_fix(ser + "_iter"))
            print("  iterations: device %d, reference order %d" % (it, its))
            assert abs(it - its) <= 0.1 * its, (it, its)
            # receivers in double: the chunked-order oracle's receivers are
            # those of this exact vector, so compare the serial ones to them
            rc, rs = _fix(chk + "_rec"), _fix(ser + "_rec")
            emax = np.abs(rs[[0, 2]]).max()
            tol = _tol(s, f)
            for ch in range(3):
                scale = emax if ch == 1 else np.abs(rs[ch]).max()
                d = np.abs(rc[ch] - rs[ch]).max() / scale
                print("  channel %d: serial vs device order %.3g of the channel scale (tol %.2g)" % (ch, d, tol))
                assert d <= 5 * tol, (f, ch, d)
            # and the float32 receivers extract_emf stores are these, rounded
            rec = ses.extract(x)
            assert np.array_equal(rec[:3], rc[:3].astype(np.complex64))
    finally:
        ses.close()


def test_layered_64_plain_fp64(workdir):
    _check_case("layered_64", MC.cases()["layered_64"], workdir)


def test_layered_64_reference_order_bitwise(workdir):
    """In the reference's serial summation order the device solve is the
    reference's: SHA-256 of the solution and the iteration count equal the
    oracle's serial-order solve (the oracle's serial order equals the
    reference build bit for bit on every golden case)."""
    from demo_matrixfree_csem_amd import _lib
    s = MC.cases()["layered_64"]
    ses = _session(s, workdir)
    L = _lib.lib()
    try:
        f = s.freqs[0]
        print("case layered_64 f=%g serial order" % f, flush=True)
        ses.set_freq(f)
        b = ses.inject()
        L.csem_set_reduction(1)
        try:
            x, it = ses.solve(b, _tol(s, f), niter=int(s.flags["niter"]))
        finally:
            L.csem_set_reduction(0)
        assert it == int(_fix("layered_64_f0_serial_iter"))
        assert hashlib.sha256(x.tobytes()).hexdigest() == str(_fix("layered_64_f0_serial_sha256"))
    finally:
        ses.close()


def test_layered_100_plain_fp64(workdir):
    _check_case("layered_100", MC.cases()["layered_100"], workdir)
