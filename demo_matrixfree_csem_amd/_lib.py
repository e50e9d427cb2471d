"""ctypes view of libcsem.so: the reference-shaped host entry points
(include/csem.h) and the device C-ABI (include/csem_dev.h).

Loading fails loudly when the library has not been built: there is no
other implementation behind these calls.
"""
import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libcsem.so")

c_cplx = C.c_double * 2
vp = C.c_void_p


class EmfT(C.Structure):
    """emf_t (include/csem.h), field for field."""
    _fields_ = [
        ("mode", C.c_int), ("verb", C.c_int), ("reciprocity", C.c_int), ("fvm", C.c_int), ("airbc", C.c_int),
        ("nfreq", C.c_int), ("freqs", C.POINTER(C.c_float)),
        ("rho_air", C.c_float), ("lextend", C.c_float), ("rho_water", C.c_float),
        ("rhomin", C.c_float), ("rhomax", C.c_float),
        ("nchsrc", C.c_int), ("nchrec", C.c_int), ("chsrc", vp), ("chrec", vp),
        ("istretch", C.c_int), ("rstretch", C.c_float),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nb", C.c_int),
        ("n1", C.c_int), ("n2", C.c_int), ("n3", C.c_int),
        ("lv", C.c_int), ("lE", C.c_int), ("lH", C.c_int), ("lEx", C.c_int), ("lEy", C.c_int), ("lEz", C.c_int),
        ("lHx", C.c_int), ("lHy", C.c_int), ("lHz", C.c_int),
        ("x1node", vp), ("x2node", vp), ("x3node", vp),
        ("x1", C.POINTER(C.c_double)), ("x2", C.POINTER(C.c_double)), ("x3", C.POINTER(C.c_double)),
        ("x1s", C.POINTER(C.c_double)), ("x2s", C.POINTER(C.c_double)), ("x3s", C.POINTER(C.c_double)),
        ("d1", C.POINTER(C.c_double)), ("d2", C.POINTER(C.c_double)), ("d3", C.POINTER(C.c_double)),
        ("d1s", C.POINTER(C.c_double)), ("d2s", C.POINTER(C.c_double)), ("d3s", C.POINTER(C.c_double)),
        ("d1min", C.c_double), ("d2min", C.c_double), ("d3min", C.c_double),
        ("x1min", C.c_double), ("x2min", C.c_double), ("x3min", C.c_double),
        ("x1max", C.c_double), ("x2max", C.c_double), ("x3max", C.c_double),
        ("n1fft", C.c_int), ("n2fft", C.c_int),
        ("d1fft", C.c_double), ("d2fft", C.c_double), ("d3fft", C.c_double),
        ("sH1kxky", C.POINTER(C.POINTER(C.c_double))), ("sH2kxky", C.POINTER(C.POINTER(C.c_double))),
        ("rho11", vp), ("rho22", vp), ("rho33", vp),
        ("sigma11", C.POINTER(C.POINTER(C.POINTER(C.c_double)))),
        ("sigma22", C.POINTER(C.POINTER(C.POINTER(C.c_double)))),
        ("sigma33", C.POINTER(C.POINTER(C.POINTER(C.c_double)))),
        ("E1", vp), ("E2", vp), ("E3", vp),
        ("dcal_fd", C.POINTER(C.POINTER(C.POINTER(C.c_float)))),
        ("ne", C.c_int), ("nh", C.c_int),
        ("xe", vp), ("xh", vp), ("be", vp), ("v", vp), ("bv", vp), ("xx", vp), ("yy", vp),
        ("I_omega_mu0", c_cplx),
        ("preco", C.c_int), ("algopt", C.c_int), ("niter", C.c_int), ("nrestart", C.c_int), ("mp", C.c_int),
        ("tol", C.c_float), ("param_extended", C.c_int),
    ]


class AcqT(C.Structure):
    _fields_ = [
        ("nsrc", C.c_int), ("nrec", C.c_int), ("nsubsrc", C.c_int), ("nsubrec", C.c_int),
        ("lensrc", C.c_float), ("lenrec", C.c_float),
        ("src_x1", C.POINTER(C.c_float)), ("src_x2", C.POINTER(C.c_float)), ("src_x3", C.POINTER(C.c_float)),
        ("src_azimuth", vp), ("src_dip", vp),
        ("rec_x1", C.POINTER(C.c_float)), ("rec_x2", C.POINTER(C.c_float)), ("rec_x3", C.POINTER(C.c_float)),
        ("rec_azimuth", vp), ("rec_dip", vp),
        ("x1min", C.c_float), ("x1max", C.c_float), ("x2min", C.c_float), ("x2max", C.c_float),
        ("x3min", C.c_float), ("x3max", C.c_float),
        ("nsrc_total", C.c_int), ("nrec_total", C.c_int), ("shot_idx", vp),
    ]


class FreqT(C.Structure):
    _fields_ = [("iomu_re", C.c_double), ("iomu_im", C.c_double),
                ("iwmu_re", C.c_double), ("iwmu_im", C.c_double)]


PRECOND_LUSGS, PRECOND_RBSGS = 0, 1  # csem_mixed_opts.precond
KRYLOV_FBICGSTAB, KRYLOV_FGMRES = 0, 1  # csem_mixed_opts.krylov


class MixedOpts(C.Structure):
    """csem_mixed_opts of include/csem_dev.h."""
    _fields_ = [("max_inner", C.c_int), ("inner_tol", C.c_double), ("max_outer", C.c_int),
                ("check_every", C.c_int), ("use_graph", C.c_int), ("precond", C.c_int),
                ("stall_iters", C.c_int), ("krylov", C.c_int), ("restart", C.c_int)]

    @classmethod
    def default(cls, **kw):
        o = cls(2000, 1e-2, 40, 8, 1, PRECOND_LUSGS, 0, KRYLOV_FBICGSTAB, 0)
        for k, v in kw.items():
            setattr(o, k, v)
        return o


# every symbol the headers in include/ declare
HOST_SYMBOLS = [
    "emf_init", "emf_free", "acq_init", "acq_free", "extend_model_init", "extend_model_free",
    "fdfd_init", "fdfd_free", "fdfd_apply", "lusgs_apply", "airbc_init", "airbc_free", "airbc_apply",
    "bicgstab", "gmres", "fbicgstab", "fgmres", "cdotprod", "inject_source", "extract_emf", "write_data",
    "find_index", "create_nugrid", "initargs", "getparint", "getparfloat", "getpardouble", "getparstring",
    "getparstringarray", "getnparint", "getnparfloat", "getnpardouble", "getnparstring", "countparname",
    "countparval", "countnparval", "alloc1", "alloc2", "alloc3", "free1", "free2", "free3", "alloc1int",
    "free1int", "alloc1float", "alloc3float", "free1float", "free3float", "alloc1double", "alloc3double",
    "free1double", "free3double", "alloc1complex", "alloc2complex", "free1complex", "free2complex",
    "alloc3complexf", "free3complexf", "err", "warn", "csem_last_status", "csem_set_exit_on_error",
    "csem_active_model", "csem_last_iterations", "csem_fdfd_prepare", "csem_model_from_emf",
    "iproc", "nproc", "ierr", "aphi_apply", "gmres_init", "gmres_free", "gmres_apply",
    "Acsr_init", "Acsr_free", "Acsr_apply", "Acsr_lusgs_apply", "Acoo_init", "Acoo_free", "Acoo_solve",
    "extract_emf_double", "write_data_double", "csem_extend_grid", "csem_model_build_device",
]
DEV_SYMBOLS = [
    "csem_last_error", "csem_freq_consts", "csem_model_create", "csem_model_create_raw", "csem_model_sigma",
    "csem_model_multiplier", "csem_model_destroy", "csem_model_len",
    "csem_model_full_len", "csem_model_bytes", "csem_apply", "csem_lusgs", "csem_curl_h", "csem_airbc",
    "csem_fft2", "csem_dot", "csem_dot_noalloc", "csem_bicgstab", "csem_fgmres", "csem_gmres", "csem_scatter_add", "csem_gather",
    "csem_malloc", "csem_free", "csem_memcpy_h2d", "csem_memcpy_d2h", "csem_memset", "csem_copy", "csem_sync",
    "csem_kernel_timing", "csem_kernel_stats", "csem_kernel_repeat", "csem_kernel_loop_stats", "csem_set_reduction",
    "csem_vec_sub", "csem_vec_divr", "csem_vec_divz", "csem_vec_subscaled_dev", "csem_vec_addscaled",
    "csem_op_apply", "csem_op_lusgs", "csem_op_aphi", "csem_op_gmres", "csem_op_host",
    "csem_gmres_prec_release", "csem_bicgstab_op", "csem_gmres_op",
    "csem_csr_create", "csem_csr_destroy", "csem_csr_nnz", "csem_csr_rows", "csem_csr_download", "csem_op_csr",
    "csem_op_csr_sgs",
    "csem_model_apply_planes", "csem_model_set_apply_planes",
    "csem_mixed_solve", "csem_apply32", "csem_lusgs32", "csem_mixed_tuning", "csem_mixed_set_tile", "csem_mixed_tuning_lu", "csem_mixed_set_lu_tiles", "csem_mixed_bytes", "csem_mixed_slot_bytes", "csem_mem_info",
]

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libcsem.so is not built (%s); run python -m demo_matrixfree_csem_amd.build"
                           % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    ip, dp, lp = C.c_int, C.c_double, C.c_long
    L.csem_last_error.restype = C.c_char_p
    L.csem_freq_consts.argtypes = [C.c_float, C.POINTER(FreqT)]
    L.csem_model_destroy.argtypes = [vp]
    for f in ("csem_model_len", "csem_model_full_len", "csem_model_bytes"):
        getattr(L, f).restype = lp
        getattr(L, f).argtypes = [vp]
    L.csem_apply.argtypes = [vp, C.POINTER(FreqT), vp, vp, vp]
    L.csem_lusgs.argtypes = [vp, C.POINTER(FreqT), vp, vp, vp]
    L.csem_curl_h.argtypes = [vp, C.POINTER(FreqT), vp, vp, vp]
    L.csem_airbc.argtypes = [vp, vp, vp, vp, vp]
    L.csem_fft2.argtypes = [vp, vp, ip, vp]
    L.csem_dot.argtypes = [lp, vp, vp, vp, vp]
    L.csem_dot_noalloc.argtypes = [lp, vp, vp, vp, vp]
    L.csem_bicgstab.argtypes = [vp, C.POINTER(FreqT), vp, vp, ip, dp, ip, vp, vp]
    L.csem_fgmres.argtypes = [vp, C.POINTER(FreqT), vp, vp, ip, dp, ip, ip, vp, vp]
    L.csem_gmres.argtypes = [vp, C.POINTER(FreqT), vp, vp, ip, dp, ip, vp, vp]
    L.csem_scatter_add.argtypes = [lp, vp, vp, vp, vp]
    L.csem_gather.argtypes = [ip, vp, vp, vp, vp, vp, vp, vp]
    L.csem_malloc.argtypes = [C.POINTER(vp), lp]
    L.csem_free.argtypes = [vp]
    L.csem_memcpy_h2d.argtypes = [vp, vp, lp, vp]
    L.csem_memcpy_d2h.argtypes = [vp, vp, lp, vp]
    L.csem_memset.argtypes = [vp, ip, lp, vp]
    L.csem_copy.argtypes = [vp, vp, lp, vp]
    L.csem_sync.argtypes = [vp]
    L.csem_kernel_timing.argtypes = [ip]
    L.csem_mixed_solve.argtypes = [vp, ip, C.POINTER(FreqT), vp, vp, vp, C.POINTER(MixedOpts), vp, vp, vp]
    L.csem_apply32.argtypes = [vp, ip, C.POINTER(FreqT), vp, vp, vp]
    L.csem_lusgs32.argtypes = [vp, ip, C.POINTER(FreqT), vp, vp, vp]
    L.csem_mixed_tuning.argtypes = [vp, ip, C.POINTER(ip), C.POINTER(ip)]
    L.csem_mixed_set_tile.argtypes = [vp, ip, ip, ip]
    L.csem_model_apply_planes.argtypes = [vp]
    L.csem_model_set_apply_planes.argtypes = [vp, ip]
    L.csem_mixed_tuning_lu.argtypes = [vp, ip, vp, C.POINTER(ip)]
    L.csem_csr_create.argtypes = [C.POINTER(vp), ip, ip, ip] + [vp] * 9 + [dp, dp]
    L.csem_csr_destroy.argtypes = [vp]
    L.csem_csr_nnz.restype = C.c_long
    L.csem_csr_nnz.argtypes = [vp]
    L.csem_csr_rows.restype = C.c_long
    L.csem_csr_rows.argtypes = [vp]
    L.csem_csr_download.argtypes = [vp, vp, vp, vp]
    L.csem_mixed_set_lu_tiles.argtypes = [vp, ip, vp, ip]
    L.csem_mixed_bytes.argtypes = [vp]
    L.csem_mixed_slot_bytes.argtypes = [vp, ip]
    L.csem_mixed_slot_bytes.restype = lp
    L.csem_mem_info.argtypes = [C.POINTER(lp), C.POINTER(lp)]
    L.csem_mixed_bytes.restype = lp
    L.csem_set_reduction.argtypes = [ip]
    L.csem_kernel_stats.argtypes = [ip, C.POINTER(C.c_char_p), C.POINTER(dp), C.POINTER(lp), C.POINTER(dp)]
    L.csem_kernel_repeat.argtypes = [ip, ip, ip]
    L.csem_kernel_loop_stats.argtypes = [ip, C.POINTER(dp), C.POINTER(lp)]
    L.csem_active_model.restype = vp
    L.csem_fdfd_prepare.argtypes = [C.POINTER(EmfT)]
    L.csem_model_from_emf.argtypes = [C.POINTER(EmfT), C.POINTER(vp)]
    L.csem_model_build_device.argtypes = [C.POINTER(AcqT), C.POINTER(EmfT), C.POINTER(vp)]
    L.csem_model_sigma.argtypes = [vp, ip, vp]
    L.csem_model_multiplier.argtypes = [vp, ip, vp]
    for f in ("emf_init", "emf_free", "fdfd_init", "fdfd_free", "extend_model_free"):
        getattr(L, f).argtypes = [C.POINTER(EmfT)]
    L.acq_init.argtypes = [C.POINTER(AcqT), C.POINTER(EmfT)]
    L.extend_model_init.argtypes = [C.POINTER(AcqT), C.POINTER(EmfT)]
    L.acq_free.argtypes = [C.POINTER(AcqT)]
    L.inject_source.argtypes = [C.POINTER(AcqT), C.POINTER(EmfT), vp]
    L.extract_emf.argtypes = [C.POINTER(AcqT), C.POINTER(EmfT), vp, ip]
    L.extract_emf_double.argtypes = [C.POINTER(AcqT), C.POINTER(EmfT), vp, vp]
    L.write_data_double.argtypes = [C.POINTER(AcqT), C.POINTER(EmfT), C.c_char_p, vp]
    L.write_data.argtypes = [C.POINTER(AcqT), C.POINTER(EmfT), C.c_char_p, vp]
    L.initargs.argtypes = [ip, C.POINTER(C.c_char_p)]
    L.fdfd_apply.argtypes = [ip, vp, vp]
    L.lusgs_apply.argtypes = [ip, vp, vp]
    L.aphi_apply.argtypes = [ip, vp, vp]
    L.gmres_init.argtypes = [ip, ip, vp]
    L.gmres_free.argtypes = []
    L.gmres_apply.argtypes = [ip, vp, vp]
    L.Acsr_init.argtypes = [C.POINTER(EmfT), ip]
    L.Acsr_apply.argtypes = [ip, vp, vp]
    L.Acsr_lusgs_apply.argtypes = [ip, vp, vp]
    L.Acsr_free.argtypes = []
    L.Acoo_init.argtypes = [C.POINTER(EmfT), ip]
    L.airbc_apply.argtypes = [C.POINTER(EmfT), vp, vp, vp]
    op_t = C.CFUNCTYPE(None, ip, vp, vp)
    L.fbicgstab.argtypes = [ip, vp, vp, vp, vp, ip, dp, ip]
    L.bicgstab.argtypes = [ip, vp, vp, vp, ip, dp, ip]
    L.fgmres.argtypes = [ip, vp, vp, vp, vp, ip, dp, ip, ip]
    L.gmres.argtypes = [ip, vp, vp, vp, ip, dp, ip, ip]
    L.cdotprod.restype = _C128
    L.cdotprod.argtypes = [ip, vp, vp]
    L.find_index.argtypes = [ip, vp, dp]
    L.create_nugrid.restype = dp
    L.create_nugrid.argtypes = [ip, dp, dp, vp, ip]
    L.alloc1complex.restype = vp
    L.alloc1complex.argtypes = [C.c_size_t]
    L.alloc3complexf.restype = C.POINTER(C.POINTER(C.POINTER(C.c_float)))
    L.alloc3complexf.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t]
    L.csem_set_exit_on_error.argtypes = [ip]
    _lib = L
    return L


class _C128(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


def last_error():
    return lib().csem_last_error().decode(errors="replace")


def check(rc, what=""):
    if rc < 0:
        raise RuntimeError("%s failed (%d): %s" % (what, rc, last_error()))
    return rc


def kernel_stats():
    """{name: (total_ms, launches, algorithmic_bytes)} since csem_kernel_timing(1)."""
    L = lib()
    out = {}
    i, n = 0, 1
    while i < n:
        name, ms, cnt, by = C.c_char_p(), C.c_double(), C.c_long(), C.c_double()
        n = L.csem_kernel_stats(i, C.byref(name), C.byref(ms), C.byref(cnt), C.byref(by))
        if n < 0:
            break
        if cnt.value:
            out[name.value.decode()] = (ms.value, cnt.value, by.value)
        i += 1
    return out


def kernel_id(name):
    """The kernel timers' id of a kernel name (as kernel_stats names it)."""
    L = lib()
    i, n = 0, 1
    while i < n:
        nm, ms, cnt, by = C.c_char_p(), C.c_double(), C.c_long(), C.c_double()
        n = L.csem_kernel_stats(i, C.byref(nm), C.byref(ms), C.byref(cnt), C.byref(by))
        if n < 0:
            break
        if nm.value.decode() == name:
            return i
        i += 1
    raise KeyError(name)


def kernel_loop_stats(name):
    """(total_ms, launches) of the repeat loops of kernel `name`."""
    ms, cnt = C.c_double(), C.c_long()
    check(lib().csem_kernel_loop_stats(kernel_id(name), C.byref(ms), C.byref(cnt)), "csem_kernel_loop_stats")
    return ms.value, cnt.value


def fn_addr(name):
    """Address of an exported function, for op_t arguments."""
    return C.cast(getattr(lib(), name), vp).value
