"""Host mirror of the reference's driver (main.c:46-145) over libcsem.so.

`Session(argv)` runs the reference's setup sequence -- initargs, emf_init,
acq_init, extend_model_init, fdfd_init -- with the same key=value flags, so
a test or a tool can then call the reference entry points (fdfd_apply,
lusgs_apply, fbicgstab, ...) on numpy host arrays, or the device C-ABI on
the model fdfd_init built.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import EmfT, AcqT, FreqT, lib, check


def freq_consts(freq):
    f = FreqT()
    lib().csem_freq_consts(C.c_float(freq), C.byref(f))
    return f


def _p(a):
    return C.c_void_p(a.ctypes.data)


class Session:
    """One extended model, as main.c builds it (one per process)."""

    def __init__(self, argv, verb=0, device=True, setup="host"):
        """The reference's setup sequence (main.c:46-92) on the flags `argv`.
        device: fdfd_init builds the active device model (else only the
        host half, csem_fdfd_prepare).  setup="device": instead of
        extend_model_init and fdfd_init, csem_model_build_device builds a
        model of this session's own on the device (`self.built`, the
        caller's to destroy), and the host keeps the 1-D tables and a copy
        of the conductivities."""
        L = lib()
        L.csem_set_exit_on_error(0)
        args = [b"csem"] + [a.encode() if isinstance(a, str) else a for a in argv]
        self._argv = (C.c_char_p * len(args))(*args)
        L.initargs(len(args), self._argv)
        self.emf, self.acq = EmfT(), AcqT()
        e = self.emf
        flags = dict(a.split("=", 1) for a in argv if "=" in a)
        # main.c:62-69 defaults
        e.mode = int(flags.get("mode", 0))
        e.verb = int(flags.get("verb", verb))
        e.niter = int(flags.get("niter", 1000))
        e.preco = int(flags.get("preco", 1))
        e.algopt = int(flags.get("algopt", 1))
        e.tol = float(np.float32(float(flags.get("tol", 1e-8))))
        e.nrestart = int(flags.get("nrestart", 10))
        e.mp = int(flags.get("mp", 30))
        self.built = C.c_void_p()
        steps = [("emf_init", lambda: L.emf_init(C.byref(e))),
                 ("acq_init", lambda: L.acq_init(C.byref(self.acq), C.byref(e)))]
        if setup == "device":
            steps.append(("csem_model_build_device",
                          lambda: L.csem_model_build_device(C.byref(self.acq), C.byref(e), C.byref(self.built))))
            device = False
        else:
            steps += [("extend_model_init", lambda: L.extend_model_init(C.byref(self.acq), C.byref(e))),
                      ("fdfd_init", (lambda: L.fdfd_init(C.byref(e))) if device
                       else (lambda: L.csem_fdfd_prepare(C.byref(e))))]
        for step, fn in steps:
            fn()
            st = L.csem_last_status()
            if st < 0:
                raise RuntimeError("%s failed (%d): %s" % (step, st, _lib.last_error()))
        self.model = L.csem_active_model() if device else None
        self.n = e.lE
        self.freqs = [e.freqs[i] for i in range(e.nfreq)]
        self._keep = []

    # ------------------------------------------------------------- facts
    @property
    def dims(self):
        e = self.emf
        return dict(n1=e.n1, n2=e.n2, n3=e.n3, nb=e.nb, nz=e.nz, lEx=e.lEx, lEy=e.lEy, lEz=e.lEz, lE=e.lE)

    def array(self, name, count):
        ptr = getattr(self.emf, name)
        return np.ctypeslib.as_array(ptr, (count,)).copy()

    def sigma(self, which):
        e = self.emf
        n1, n2, n3 = e.n1, e.n2, e.nz + 2 * e.nb
        cnt = {11: n1 * (n2 + 1) * (n3 + 1), 22: (n1 + 1) * n2 * (n3 + 1), 33: (n1 + 1) * (n2 + 1) * n3}[which]
        p = getattr(e, "sigma%d" % which)[0][0]
        return np.ctypeslib.as_array(p, (cnt,)).copy()

    # ------------------------------------------------------- frequency
    def set_freq(self, freq):
        """emf->I_omega_mu0 = I*2*PI*f*mu0 (main.c:95)."""
        f = freq_consts(freq)
        self.emf.I_omega_mu0[0] = f.iomu_re
        self.emf.I_omega_mu0[1] = f.iomu_im
        self.freq = freq
        return f

    # -------------------------------------------------- host entry points
    def fdfd_apply(self, x):
        x = np.ascontiguousarray(x, np.complex128)
        y = np.zeros(self.n, np.complex128)
        lib().fdfd_apply(self.n, _p(x), _p(y))
        self._status("fdfd_apply")
        return y

    def lusgs_apply(self, x):
        x = np.ascontiguousarray(x, np.complex128)
        y = np.zeros(self.n, np.complex128)
        lib().lusgs_apply(self.n, _p(x), _p(y))
        self._status("lusgs_apply")
        return y

    def airbc_apply(self, hz):
        e = self.emf
        hz = np.ascontiguousarray(hz, np.complex128)
        hx = np.zeros((e.n1 + 1) * e.n2, np.complex128)
        hy = np.zeros(e.n1 * (e.n2 + 1), np.complex128)
        lib().airbc_apply(C.byref(e), _p(hz), _p(hx), _p(hy))
        self._status("airbc_apply")
        return hx, hy

    def cdotprod(self, a, b):
        a = np.ascontiguousarray(a, np.complex128)
        b = np.ascontiguousarray(b, np.complex128)
        r = lib().cdotprod(len(a), _p(a), _p(b))
        self._status("cdotprod")
        return complex(r.re, r.im)

    def inject(self):
        b = np.zeros(self.n, np.complex128)
        lib().inject_source(C.byref(self.acq), C.byref(self.emf), _p(b))
        return b

    def extract(self, x, ifreq=0):
        """extract_emf into a fresh dcal_fd; returns complex64 (nchrec, nrec)."""
        e, a = self.emf, self.acq
        L = lib()
        if not e.dcal_fd:
            e.dcal_fd = L.alloc3complexf(a.nrec, e.nfreq, e.nchrec)
        base = C.cast(e.dcal_fd[0][0], C.c_void_p).value
        total = a.nrec * e.nfreq * e.nchrec
        C.memset(base, 0, total * 8)
        x = np.ascontiguousarray(x, np.complex128)
        L.extract_emf(C.byref(a), C.byref(e), _p(x), ifreq)
        arr = np.ctypeslib.as_array(C.cast(base, C.POINTER(C.c_float)), (total * 2,)).copy().view(np.complex64)
        return arr.reshape(e.nchrec, e.nfreq, a.nrec)[:, ifreq, :]

    def extract_double(self, x):
        """extract_emf_double: every receiver channel in double precision
        (the values extract_emf rounds to float); complex128 (nchrec, nrec)."""
        e, a = self.emf, self.acq
        x = np.ascontiguousarray(x, np.complex128)
        out = np.zeros(e.nchrec * a.nrec, np.complex128)
        lib().extract_emf_double(C.byref(a), C.byref(e), _p(x), _p(out))
        return out.reshape(e.nchrec, a.nrec)

    def inject_at(self, sx, sy, sz):
        """inject_source for a source at (sx, sy, sz) (the first source's
        slot of acq_t, acq_init_free.c:65-69, is repointed for the call)."""
        a = self.acq
        old = (a.src_x1[0], a.src_x2[0], a.src_x3[0])
        a.src_x1[0], a.src_x2[0], a.src_x3[0] = sx, sy, sz
        try:
            return self.inject()
        finally:
            a.src_x1[0], a.src_x2[0], a.src_x3[0] = old

    def extract_at(self, x, receivers, double=False):
        """extract_emf at the given receiver positions [(x, y, z), ...];
        returns complex64 (nchrec, nrec), or with `double` the values
        before the float store (extract_emf_double), complex128."""
        a, e, L = self.acq, self.emf, lib()
        nr = len(receivers)
        r = np.asarray(receivers, np.float32).reshape(nr, -1)
        # a pointer read from a ctypes Structure aliases the field: keep the
        # addresses, not the pointer objects
        names = ("rec_x1", "rec_x2", "rec_x3")
        keep = [C.cast(getattr(a, k), C.c_void_p).value for k in names]
        keep_fd = C.cast(e.dcal_fd, C.c_void_p).value
        nrec = a.nrec
        bufs = [np.ascontiguousarray(r[:, c]) for c in range(3)]
        try:
            a.nrec = nr
            for k, b in zip(names, bufs):
                setattr(a, k, C.cast(b.ctypes.data, C.POINTER(C.c_float)))
            if double:
                return self.extract_double(x)
            e.dcal_fd = L.alloc3complexf(nr, e.nfreq, e.nchrec)
            out = self.extract(x)
            L.free3complexf(e.dcal_fd)
            return out
        finally:
            a.nrec = nrec
            for k, v in zip(names, keep):
                setattr(a, k, C.cast(C.c_void_p(v), C.POINTER(C.c_float)))
            e.dcal_fd = C.cast(C.c_void_p(keep_fd), type(e.dcal_fd))

    def solve(self, b, tol, niter=None, algopt=None, preco=None, nrestart=None, x0=None, verb=0, mp=None):
        """The solver call of main.c:104-118; returns (x, iterations).
        preco 0 none, 1 LU-SGS, 2 inner GMRES(mp) on fdfd_apply, 3 A-phi;
        algopt 1 (F-)BiCGStab, 2 (F)GMRES(nrestart).  With verb=1 the
        solver writes its iterate_*.txt history to the current directory,
        as the reference's does (solver.c:127-141)."""
        e = self.emf
        L = lib()
        niter = e.niter if niter is None else niter
        algopt = e.algopt if algopt is None else algopt
        preco = e.preco if preco is None else preco
        nrestart = e.nrestart if nrestart is None else nrestart
        mp = e.mp if mp is None else mp
        x = np.zeros(self.n, np.complex128) if x0 is None else np.array(x0, np.complex128)
        b = np.ascontiguousarray(b, np.complex128)
        A = _lib.fn_addr("fdfd_apply")
        M = _lib.fn_addr({1: "lusgs_apply", 2: "gmres_apply", 3: "aphi_apply"}.get(preco, "lusgs_apply"))
        if preco == 2:
            L.gmres_init(self.n, mp, A)
            self._status("gmres_init")
        try:
            if preco == 0 and algopt == 1:
                L.bicgstab(self.n, _p(x), _p(b), A, niter, tol, verb)
            elif preco == 0:
                L.gmres(self.n, _p(x), _p(b), A, niter // nrestart, tol, nrestart, verb)
            elif algopt == 1:
                L.fbicgstab(self.n, _p(x), _p(b), A, M, niter, tol, verb)
            else:
                L.fgmres(self.n, _p(x), _p(b), A, M, niter // nrestart, tol, nrestart, verb)
            self._status("solve")
        finally:
            if preco == 2:
                L.gmres_free()
        return x, L.csem_last_iterations()

    def aphi_apply(self, x):
        x = np.ascontiguousarray(x, np.complex128)
        y = np.zeros(self.n, np.complex128)
        lib().aphi_apply(self.n, _p(x), _p(y))
        self._status("aphi_apply")
        return y

    def gmres_apply(self, b, mp):
        """gmres_init(n, mp, fdfd_apply); gmres_apply(b -> x); gmres_free"""
        L = lib()
        b = np.ascontiguousarray(b, np.complex128)
        x = np.zeros(self.n, np.complex128)
        L.gmres_init(self.n, mp, _lib.fn_addr("fdfd_apply"))
        self._status("gmres_init")
        try:
            L.gmres_apply(self.n, _p(b), _p(x))
            self._status("gmres_apply")
        finally:
            L.gmres_free()
        return x

    def solve_mixed(self, bs, freqs, tols, opts=None):
        """Batched mixed-precision solve (csem_mixed_solve) of the right-hand
        sides bs (list of complex128 vectors), RHS r at frequency freqs[r]
        to true relative residual tols[r].  Returns (xs, iters, relres,
        converged count)."""
        L = lib()
        R, n = len(bs), self.n
        fr = (_lib.FreqT * R)(*[freq_consts(f) for f in freqs])
        db, dx = C.c_void_p(), C.c_void_p()
        _lib.check(L.csem_malloc(C.byref(db), R * n * 16), "csem_malloc")
        _lib.check(L.csem_malloc(C.byref(dx), R * n * 16), "csem_malloc")
        try:
            hb = np.ascontiguousarray(np.concatenate(bs), np.complex128)
            _lib.check(L.csem_memcpy_h2d(db, _p(hb), R * n * 16, None), "upload")
            tol = np.ascontiguousarray(tols, np.float64)
            its = np.zeros(R, np.int32)
            rel = np.zeros(R)
            o = opts if opts is not None else _lib.MixedOpts.default()
            nconv = L.csem_mixed_solve(self.model, R, fr, dx, db, _p(tol), C.byref(o), _p(its), _p(rel), None)
            _lib.check(nconv, "csem_mixed_solve")
            hx = np.zeros(R * n, np.complex128)
            _lib.check(L.csem_memcpy_d2h(_p(hx), dx, R * n * 16, None), "download")
            _lib.check(L.csem_sync(None), "sync")
        finally:
            L.csem_free(db)
            L.csem_free(dx)
        return [hx[r * n:(r + 1) * n].copy() for r in range(R)], its, rel, nconv

    def _status(self, what):
        st = lib().csem_last_status()
        if st < 0:
            raise RuntimeError("%s failed (%d): %s" % (what, st, _lib.last_error()))

    def close(self):
        if self.model:
            lib().fdfd_free(C.byref(self.emf))
            self.model = None

    def multipliers(self):
        """The airwave multipliers sH1, sH2 airbc_init computed (nfft x nfft)."""
        e = self.emf
        N = e.n1fft
        out = []
        for p in (e.sH1kxky, e.sH2kxky):
            a = np.ctypeslib.as_array(p[0], (2 * N * N,)).copy().view(np.complex128)
            out.append(a)
        return out
