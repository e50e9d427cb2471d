"""The matrix-storage path (Acsr.c): the operator as a CSR matrix.

Pins, on the CPU (the oracle's restatement, oracle/csem_oracle.c
orc_csr_*):
  * against the reference build's own Acsr_init arrays
    (tests/golden/ref_acsr.npz): the column indices and values are the
    same bit for bit; the row pointers are the reference's plus one from
    row 1 on -- the reference writes row_ptr[r + 1] before it stores row
    r's diagonal (Acsr.c:425-428), so its ranges put every diagonal in the
    next row; the oracle's ranges hold each row's own;
  * against the reference's own products: the oracle's product and
    Gauss-Seidel loops, run on the reference's row pointers, reproduce the
    reference's Acsr_apply and Acsr_lusgs_apply outputs bit for bit (the
    same loops and arithmetic, so the oracle's are the reference's);
  * by mathematics: every row that fdfd_apply does not zero equals
    i w mu0 times fdfd_apply's row (to 4e-15 of the row's largest term),
    and the symmetric Gauss-Seidel equals the two triangular solves
    (D + L) y1 = x, (D + U) y = x - L y1 (scipy, 1e-13).
On the GPU (through the drop-in Acsr_init / Acsr_apply /
Acsr_lusgs_apply and the device C-ABI): the device matrix, product and
Gauss-Seidel equal the oracle's bit for bit, and F-BiCGStab over
(Acsr_apply, Acsr_lusgs_apply) runs device-resident and reaches tol in
the true residual of the assembled matrix.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
import oracle as O  # noqa: E402

from demo_matrixfree_csem_amd import surveys  # noqa: E402

GOLD = np.load(os.path.join(HERE, "golden", "ref_acsr.npz"))
# the cases of tests/golden/make_golden.py ACSR_CASES
CASES = {"c0": (dict(seed=8, nx=5, ny=4, nz=3, airbc=0), dict(nb=2, rho_air=3.0), 0.75),
         "c1": (dict(seed=9, nx=4, ny=5, nz=4, airbc=0), dict(nb=2, rho_air=10.0), 1.25)}


def _survey(name):
    kw, fl, f = CASES[name]
    s = surveys.tiny(**kw)
    s.flags.update(fl)
    s.freqs = [f]
    return s


def _oracle(s):
    m = s.model
    orc = O.Oracle(m.x1node, m.x2node, m.x3node, m.rho11, m.rho22, m.rho33, nb=s.flags["nb"], airbc=0,
                   rho_air=s.flags["rho_air"], nfft=64)
    orc.set_freq(s.freqs[0])
    return orc


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_csr_is_the_reference_assembly(name):
    orc = _oracle(_survey(name))
    rp, col, val = orc.csr()
    g_rp, g_col, g_val = GOLD[name + "_row_ptr"], GOLD[name + "_col"], GOLD[name + "_val"]
    assert np.array_equal(col, g_col)
    assert np.array_equal(val.view(np.float64), g_val.view(np.float64))
    # the reference's range of row r ends before r's diagonal
    assert rp[0] == g_rp[0] == 0 and np.array_equal(rp[1:], g_rp[1:] + 1)
    # every row's own range ends with its diagonal
    assert np.array_equal(col[rp[1:] - 1], np.arange(len(rp) - 1))


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_products_are_the_reference_loops(name):
    orc = _oracle(_survey(name))
    _, col, val = orc.csr()
    x, g_rp = GOLD[name + "_x"], GOLD[name + "_row_ptr"]
    y = O.Oracle.csr_apply(g_rp, col, val, x)
    assert np.array_equal(y.view(np.float64), GOLD[name + "_apply"].view(np.float64))
    ys = O.Oracle.csr_sgs(g_rp, col, val, x)
    assert np.array_equal(ys.view(np.float64), GOLD[name + "_sgs"].view(np.float64))


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_csr_rows_are_iomu_times_fdfd(name):
    s = _survey(name)
    orc = _oracle(s)
    rp, col, val = orc.csr()
    x = GOLD[name + "_x"]
    y = O.Oracle.csr_apply(rp, col, val, x)
    ya = orc.apply(x)
    iomu = complex(0.0, 2.0 * np.pi * float(np.float32(s.freqs[0])) * 4.0 * np.pi * 1e-7)
    rows = np.nonzero(np.abs(ya) > 0)[0]
    assert len(rows) > len(ya) // 2
    big = np.array([np.abs(val[rp[r]:rp[r + 1]] * x[col[rp[r]:rp[r + 1]]]).max() for r in rows])
    err = np.abs(y[rows] - iomu * ya[rows]) / big
    assert err.max() < 4e-15, err.max()
    # the rows fdfd_apply zeroes are the boundary rows with truncated stencils
    assert np.all(np.diff(rp)[np.abs(ya) == 0] < 13)


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_sgs_is_two_triangular_solves(name):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    orc = _oracle(_survey(name))
    rp, col, val = orc.csr()
    n = len(rp) - 1
    A = sp.csr_matrix((val, col, rp), shape=(n, n))
    x = GOLD[name + "_x"]
    Lo, Up = sp.tril(A, format="csr"), sp.triu(A, format="csr")
    y1 = spl.spsolve_triangular(Lo, x, lower=True)
    y = spl.spsolve_triangular(Up, x - (Lo - sp.diags(A.diagonal())) @ y1, lower=False)
    ys = O.Oracle.csr_sgs(rp, col, val, x)
    assert np.abs(ys - y).max() < 1e-13 * np.abs(y).max()


def _session(s, workdir):
    from demo_matrixfree_csem_amd.session import Session
    surveys.write_files(workdir, s)
    return Session(s.argv(workdir))


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_acsr_device_equals_oracle(workdir, name):
    from demo_matrixfree_csem_amd import _lib
    print("case acsr %s" % name, flush=True)
    s = _survey(name)
    orc = _oracle(s)
    rp, col, val = orc.csr()
    ses = _session(s, workdir)
    L = _lib.lib()
    try:
        L.Acsr_init(C.byref(ses.emf), 0)
        _lib.check(L.csem_last_status(), "Acsr_init")
        n = len(rp) - 1
        assert ses.n == n
        x = np.ascontiguousarray(GOLD[name + "_x"])
        y = np.zeros(n, np.complex128)
        L.Acsr_apply(n, x.ctypes.data, y.ctypes.data)
        _lib.check(L.csem_last_status(), "Acsr_apply")
        assert np.array_equal(y.view(np.float64), O.Oracle.csr_apply(rp, col, val, x).view(np.float64))
        L.Acsr_lusgs_apply(n, x.ctypes.data, y.ctypes.data)
        _lib.check(L.csem_last_status(), "Acsr_lusgs_apply")
        assert np.array_equal(y.view(np.float64), O.Oracle.csr_sgs(rp, col, val, x).view(np.float64))
        L.Acsr_free()
        # the device C-ABI: the assembled arrays themselves
        e = ses.emf
        arr = lambda p, k: np.ctypeslib.as_array(p, shape=(k,))
        n1, n2, n3 = e.n1, e.n2, e.n3
        A = C.c_void_p()
        lEx, lEy = n1 * (n2 + 1) * (n3 + 1), (n1 + 1) * n2 * (n3 + 1)
        iomu = 2.0 * np.pi * float(np.float32(s.freqs[0])) * (4.0 * np.pi * 1e-7)  # I*2.*PI*f*mu0, main.c:95
        vp = lambda p: C.cast(p, C.c_void_p)
        _lib.check(L.csem_csr_create(C.byref(A), n1, n2, n3, vp(e.d1), vp(e.d2), vp(e.d3), vp(e.d1s), vp(e.d2s),
                                     vp(e.d3s),
                                     arr(e.sigma11[0][0], lEx).ctypes.data, arr(e.sigma22[0][0], lEy).ctypes.data,
                                     arr(e.sigma33[0][0], n - lEx - lEy).ctypes.data, C.c_double(0.0),
                                     C.c_double(iomu)), "csem_csr_create")
        try:
            assert L.csem_csr_nnz(A) == len(col) and L.csem_csr_rows(A) == n
            drp, dcol = np.zeros_like(rp), np.zeros_like(col)
            dval = np.zeros_like(val)
            _lib.check(L.csem_csr_download(A, drp.ctypes.data, dcol.ctypes.data, dval.ctypes.data), "download")
            assert np.array_equal(drp, rp) and np.array_equal(dcol, col)
            assert np.array_equal(dval.view(np.float64), val.view(np.float64))
        finally:
            L.csem_csr_destroy(A)
    finally:
        ses.close()


@pytest.mark.gpu
def test_acsr_solve_device_resident(workdir):
    """main.c's main2 (preco=1, algopt=1): F-BiCGStab over (Acsr_apply,
    Acsr_lusgs_apply), mapped to the device matrix by the driver, reaches
    tol in the true residual |b - A x| / |b| of the assembled matrix."""
    import scipy.sparse as sp
    from demo_matrixfree_csem_amd import _lib
    print("case acsr solve", flush=True)
    s = surveys.tiny(10, nx=8, ny=8, nz=6, airbc=0)
    s.flags.update(nb=3, rho_air=10.0)
    s.freqs = [0.5]
    orc = _oracle(s)
    rp, col, val = orc.csr()
    ses = _session(s, workdir)
    L = _lib.lib()
    try:
        L.Acsr_init(C.byref(ses.emf), 0)
        _lib.check(L.csem_last_status(), "Acsr_init")
        ses.set_freq(0.5)
        b = np.ascontiguousarray(ses.inject())
        x = np.zeros_like(b)
        tol = 1e-8 / 0.5
        L.fbicgstab(ses.n, x.ctypes.data, b.ctypes.data, _lib.fn_addr("Acsr_apply"), _lib.fn_addr("Acsr_lusgs_apply"),
                    2000, C.c_double(tol), 0)
        _lib.check(L.csem_last_status(), "fbicgstab")
        it = L.csem_last_iterations()
        import scipy.sparse.linalg as spl
        A = sp.csr_matrix((val, col, rp), shape=(ses.n, ses.n))
        rel = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        xd = spl.spsolve(A.tocsc(), b)
        err = np.linalg.norm(x - xd) / np.linalg.norm(xd)
        print("converged at %d iterations, true relres %.2e, vs a direct solve %.2e" % (it, rel, err), flush=True)
        assert 0 < it < 2000 and rel < 10 * tol and err < 1e-5
        L.Acsr_free()
    finally:
        ses.close()
