"""ctypes binding of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The product package never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

RED_SERIAL, RED_CHUNKED = 0, 1


class _C128(C.Structure):
    """double _Complex as returned by value (two SSE registers)."""
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, ip, dp, fp = C.c_void_p, C.c_int, C.c_double, C.c_float
        L.orc_build.restype = vp
        L.orc_build.argtypes = [ip, ip, ip, vp, vp, vp, vp, vp, vp, ip, fp, ip, fp, ip, ip, ip]
        L.orc_free.argtypes = [vp]
        L.orc_info.restype = C.c_long
        L.orc_info.argtypes = [vp, ip]
        L.orc_finfo.restype = dp
        L.orc_finfo.argtypes = [vp, ip]
        L.orc_array.restype = vp
        L.orc_array.argtypes = [vp, C.c_char_p]
        L.orc_set_red.argtypes = [vp, ip]
        L.orc_set_mp.argtypes = [vp, ip]
        L.orc_set_freq.argtypes = [vp, fp]
        for f in ("orc_fdfd_apply", "orc_lusgs_apply", "orc_aphi_apply", "orc_gmres_prec"):
            getattr(L, f).argtypes = [vp, vp, vp]
        L.orc_airbc_apply.argtypes = [vp, vp, vp, vp]
        L.orc_airbc_stage.argtypes = [vp, ip, vp, vp, vp]
        L.orc_fft2.argtypes = [vp, vp, ip]
        L.orc_dot.restype = _C128
        L.orc_dot.argtypes = [vp, C.c_long, vp, vp]
        L.orc_fbicgstab.restype = ip
        L.orc_fbicgstab.argtypes = [vp, vp, vp, ip, dp, ip, vp]
        L.orc_fgmres.restype = ip
        L.orc_fgmres.argtypes = [vp, vp, vp, ip, dp, ip, ip, vp]
        L.orc_gmres.restype = ip
        L.orc_gmres.argtypes = [vp, vp, vp, ip, dp, ip, vp]
        L.orc_inject.argtypes = [vp, ip, fp, fp, fp, vp]
        L.orc_extract.restype = _C128
        L.orc_extract.argtypes = [vp, ip, fp, fp, fp, vp]
        L.orc_create_nugrid.restype = dp
        L.orc_create_nugrid.argtypes = [ip, dp, dp, vp, ip]
        L.orc_find_index.restype = ip
        L.orc_find_index.argtypes = [ip, vp, dp]
        L.orc_csr_build.restype = C.c_long
        L.orc_csr_build.argtypes = [vp, vp, vp, vp]
        L.orc_csr_apply.argtypes = [C.c_long, vp, vp, vp, vp, vp]
        L.orc_csr_sgs.argtypes = [C.c_long, vp, vp, vp, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """One extended model plus the per-frequency operator state."""

    def __init__(self, x1node, x2node, x3node, rho11, rho22, rho33, nb=10,
                 lextend=-1.0, istretch=1, rho_air=1e6, airbc=0, fvm=0, nfft=1024):
        L = lib()
        self._keep = [np.ascontiguousarray(a, np.float32) for a in
                      (x1node, x2node, x3node, rho11, rho22, rho33)]
        nx, ny, nz = len(x1node) - 1, len(x2node) - 1, len(x3node) - 1
        self.h = L.orc_build(nx, ny, nz, *[_p(a) for a in self._keep], nb, lextend,
                             istretch, rho_air, airbc, fvm, nfft)
        self.airbc, self.fvm, self.nb = airbc, fvm, nb
        self.nx, self.ny, self.nz = nx, ny, nz
        i = lambda w: L.orc_info(self.h, w)
        self.n1, self.n2, self.n3, self.n3f = i(0), i(1), i(2), i(3)
        self.lEx, self.lEy, self.lEz, self.lE = i(4), i(5), i(6), i(7)
        self.fEx, self.fEy, self.fEz, self.fE = i(8), i(9), i(10), i(11)
        self.nfft = i(12)
        self.rho_water = L.orc_finfo(self.h, 0)
        self.lextend = L.orc_finfo(self.h, 1)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    def array(self, name):
        n1, n2, n3 = self.n1, self.n2, self.n3f
        sizes = {"x1": n1 + 1, "x2": n2 + 1, "x3": n3 + 1, "sigma11": n1 * (n2 + 1) * (n3 + 1),
                 "sigma22": (n1 + 1) * n2 * (n3 + 1), "sigma33": (n1 + 1) * (n2 + 1) * n3,
                "cell11": n1 * n2 * n3, "cell22": n1 * n2 * n3, "cell33": n1 * n2 * n3}
        for a in ("x", "d"):
            for ax, n in ((1, n1), (2, n2), (3, n3)):
                sizes["%s%d" % (a, ax)] = n + 1
                sizes["%s%ds" % (a, ax)] = n + 1
        cplx = {"sH1": self.nfft ** 2, "sH2": self.nfft ** 2,
                "Hxg": (n1 + 1) * n2, "Hyg": n1 * (n2 + 1)}
        ints = {"jx": self.nfft, "jy": self.nfft}
        dbl = {"wx": self.nfft, "wy": self.nfft}
        ptr = lib().orc_array(self.h, name.encode())
        if name in sizes:
            t, n = C.c_double, sizes[name]
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(t)), (n,)).copy()
        if name in cplx:
            n = cplx[name]
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_double)), (2 * n,)).copy().view(np.complex128)
        if name in ints:
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_int)), (ints[name],)).copy()
        if name in dbl:
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_double)), (dbl[name],)).copy()
        raise KeyError(name)

    def h_field(self):
        """H = curl E / (i w mu0) of the last apply's input (fdfd.c:125-149)
        as (Hx, Hy, Hz): Hx (n1+1)*n2*n3, Hy n1*(n2+1)*n3, Hz n1*n2*(n3+1)
        on the operator grid."""
        L = lib()
        n1, n2, n3 = self.n1, self.n2, self.n3
        lhx, lhz, off = L.orc_info(self.h, 13), L.orc_info(self.h, 15), L.orc_info(self.h, 16)
        lhy = n1 * (n2 + 1) * n3
        ptr = C.cast(L.orc_array(self.h, b"xh"), C.POINTER(C.c_double))
        xh = np.ctypeslib.as_array(ptr, (2 * (off + lhz),)).copy().view(np.complex128)
        return xh[:lhx], xh[lhx:lhx + lhy], xh[off:off + lhz]

    def set_red(self, red):
        lib().orc_set_red(self.h, red)

    def set_freq(self, f):
        lib().orc_set_freq(self.h, f)

    def apply(self, x):
        x = np.ascontiguousarray(x, np.complex128)
        y = np.zeros(self.lE, np.complex128)
        lib().orc_fdfd_apply(self.h, _p(x), _p(y))
        return y

    def aphi(self, x):
        x = np.ascontiguousarray(x, np.complex128)
        y = np.zeros(self.lE, np.complex128)
        lib().orc_aphi_apply(self.h, _p(x), _p(y))
        return y

    def gmres_prec(self, b, mp=None):
        if mp is not None:
            lib().orc_set_mp(self.h, mp)
        b = np.ascontiguousarray(b, np.complex128)
        x = np.zeros(self.lE, np.complex128)
        lib().orc_gmres_prec(self.h, _p(b), _p(x))
        return x

    def set_mp(self, mp):
        lib().orc_set_mp(self.h, mp)

    def lusgs(self, x):
        x = np.ascontiguousarray(x, np.complex128)
        y = np.zeros(self.lE, np.complex128)
        lib().orc_lusgs_apply(self.h, _p(x), _p(y))
        return y

    def airbc_apply(self, hz):
        hz = np.ascontiguousarray(hz, np.complex128)
        hx = np.zeros((self.n1 + 1) * self.n2, np.complex128)
        hy = np.zeros(self.n1 * (self.n2 + 1), np.complex128)
        lib().orc_airbc_apply(self.h, _p(hz), _p(hx), _p(hy))
        return hx, hy

    def airbc_stage(self, stage, data):
        """One stage of airbc_apply (see orc_airbc_stage): 0 forward map
        (returns N*N), 1 spectral continuation (returns Hx, Hy on the
        uniform grid), 2 / 3 back maps to ghost Hx / Hy."""
        N, n1, n2 = self.nfft, self.n1, self.n2
        d = np.ascontiguousarray(data, np.complex128)
        size = {0: N * N, 1: N * N, 2: (n1 + 1) * n2, 3: n1 * (n2 + 1)}[stage]
        out = np.zeros(size, np.complex128)
        out2 = np.zeros(size if stage == 1 else 1, np.complex128)
        lib().orc_airbc_stage(self.h, stage, _p(d), _p(out), _p(out2))
        return (out, out2) if stage == 1 else out

    def fft2(self, data, direction):
        d = np.ascontiguousarray(data, np.complex128).copy()
        lib().orc_fft2(self.h, _p(d), direction)
        return d

    def dot(self, a, b):
        a = np.ascontiguousarray(a, np.complex128)
        b = np.ascontiguousarray(b, np.complex128)
        r = lib().orc_dot(self.h, len(a), _p(a), _p(b))
        return complex(r.re, r.im)

    def inject(self, ch, sx, sy, sz, b=None):
        if b is None:
            b = np.zeros(self.lE, np.complex128)
        lib().orc_inject(self.h, ch, sx, sy, sz, _p(b))
        return b

    def extract(self, ch, rx, ry, rz, x):
        x = np.ascontiguousarray(x, np.complex128)
        r = lib().orc_extract(self.h, ch, rx, ry, rz, _p(x))
        return complex(r.re, r.im)

    def fbicgstab(self, b, niter, tol, precond=1, x=None):
        x = np.zeros(self.lE, np.complex128) if x is None else x.copy()
        b = np.ascontiguousarray(b, np.complex128)
        hist = np.zeros(niter + 2)
        it = lib().orc_fbicgstab(self.h, _p(x), _p(b), niter, tol, precond, _p(hist))
        return x, it, hist[:it + 2]

    def fgmres(self, b, nouter, tol, m, precond=1, x=None):
        x = np.zeros(self.lE, np.complex128) if x is None else x.copy()
        b = np.ascontiguousarray(b, np.complex128)
        hist = np.zeros(nouter * m + 2)
        it = lib().orc_fgmres(self.h, _p(x), _p(b), nouter, tol, m, precond, _p(hist))
        return x, it, hist[:it + 1]

    def gmres(self, b, nouter, tol, m, x=None):
        x = np.zeros(self.lE, np.complex128) if x is None else x.copy()
        b = np.ascontiguousarray(b, np.complex128)
        hist = np.zeros(nouter * m + 2)
        it = lib().orc_gmres(self.h, _p(x), _p(b), nouter, tol, m, _p(hist))
        return x, it, hist[:it + 1]

    def csr(self):
        """Acsr_init's matrix on the full grid at the current frequency:
        (row_ptr, col, val), each row's range holding its own diagonal."""
        L = lib()
        nnz = L.orc_csr_build(self.h, None, None, None)
        rows = int(self.fE)
        rp = np.zeros(rows + 1, np.int32)
        col = np.zeros(nnz, np.int32)
        val = np.zeros(nnz, np.complex128)
        L.orc_csr_build(self.h, _p(rp), _p(col), _p(val))
        return rp, col, val

    @staticmethod
    def csr_apply(rp, col, val, x):
        y = np.zeros(len(rp) - 1, np.complex128)
        lib().orc_csr_apply(len(rp) - 1, _p(rp), _p(col), _p(val), _p(np.ascontiguousarray(x, np.complex128)), _p(y))
        return y

    @staticmethod
    def csr_sgs(rp, col, val, x):
        y = np.zeros(len(rp) - 1, np.complex128)
        lib().orc_csr_sgs(len(rp) - 1, _p(rp), _p(col), _p(val), _p(np.ascontiguousarray(x, np.complex128)), _p(y))
        return y


def run_ref(workdir, flags, mode, hfreq=None, hin=None, hout="ref", timeout=600):
    """Run the reference harness in workdir; returns its stdout."""
    args = [REF_HARNESS] + list(flags) + ["hmode=" + mode, "hout=" + hout]
    if hfreq is not None:
        args.append("hfreq=%r" % float(hfreq))
    if hin is not None:
        args.append("hin=" + hin)
    r = subprocess.run(args, cwd=workdir, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError("ref_harness failed (%d): %s %s" % (r.returncode, r.stdout[-2000:], r.stderr[-2000:]))
    return r.stdout
