"""csem_dot_noalloc (the allocation-free dot product) against csem_dot,
through the C-ABI, on device buffers: bit-identical results on two calls
in a row, the second with a smaller input (the shared scratch keeps the
first call's partial sums, and every partial the second reads is
rewritten), plus an empty input and the size limit."""
import ctypes as C

import numpy as np
import pytest

from demo_matrixfree_csem_amd import _lib

pytestmark = pytest.mark.gpu

MAX_N = 16384 * 4096


def _dev(a):
    L = _lib.lib()
    p = C.c_void_p()
    assert L.csem_malloc(C.byref(p), max(a.nbytes, 16)) == 0
    if a.nbytes:
        assert L.csem_memcpy_h2d(p, a.ctypes.data, a.nbytes, None) == 0
    return p


def _dot(fn, n, da, db, dout):
    L = _lib.lib()
    out = np.zeros(2)
    assert fn(n, da, db, dout, None) == 0, _lib.last_error()
    assert L.csem_memcpy_d2h(out.ctypes.data, dout, 16, None) == 0
    assert L.csem_sync(None) == 0
    return out


def test_dot_noalloc_matches_dot():
    L = _lib.lib()
    r = np.random.default_rng(3)
    n1 = 3_000_017
    a = r.standard_normal(n1) + 1j * r.standard_normal(n1)
    b = r.standard_normal(n1) + 1j * r.standard_normal(n1)
    da, db, dout = _dev(a), _dev(b), _dev(np.zeros(4))
    try:
        print("case dot_noalloc n=%d then n=5003, 0" % n1, flush=True)
        for n in (n1, 5003, 0):
            ref = _dot(L.csem_dot, n, da, db, dout)
            got = _dot(L.csem_dot_noalloc, n, da, db, dout)
            assert np.array_equal(got, ref), (n, got, ref)
            exact = np.vdot(b[:n], a[:n])  # sum a conj(b)
            assert abs(complex(*got) - exact) <= 1e-10 * max(1.0, abs(exact)) * np.sqrt(max(n, 1))
        # beyond the scratch: rejected before any launch
        assert L.csem_dot_noalloc(MAX_N + 1, da, db, dout, None) < 0
        assert b"outside" in L.csem_last_error()
    finally:
        for p in (da, db, dout):
            L.csem_free(p)
