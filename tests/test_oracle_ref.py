"""Pin the oracle (oracle/csem_oracle.c).

1. Against the reference's own routines: golden vectors written by
   tests/golden/make_golden.py from the reference sources compiled as-is
   (oracle/Makefile target `ref`), and -- when that build is present -- live
   runs on more seeds.  Everything must agree bit for bit.
2. Against the reference's checked-in run_layered data (extended grid and
   model dumps).
3. Against mathematics for what no reference build can run here (the
   airwave transform needs FFTW, which is absent): the 2-D DFT against
   numpy.fft, the multipliers against their closed form, LU-SGS against a
   dense (D+U)^-1 D (D+L)^-1 solve, curl(grad) = 0, the complex symmetry of
   the volume-scaled operator, and a converged solve against a dense direct
   solve.
"""
import os

import numpy as np
import pytest

from demo_matrixfree_csem_amd import surveys
import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
MU0 = 4e-7 * np.pi


def _rand(n, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal(n) + 1j * r.standard_normal(n)


def _oracle(s, **kw):
    m = s.model
    f = s.flags
    return O.Oracle(m.x1node, m.x2node, m.x3node, m.rho11, m.rho22, m.rho33, nb=f["nb"],
                    airbc=f["airbc"], fvm=f.get("fvm", 0), rho_air=float(f.get("rho_air", 1e6)), **kw)


# ---------------------------------------------------------------- golden

def _cases():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg.CASES


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(GOLDEN, "ref_tiny.npz"))


@pytest.mark.parametrize("case", [c[0] for c in _cases()])
def test_golden_grid_and_operators(golden, case):
    name, kw, fl = [c for c in _cases() if c[0] == case][0]
    s = surveys.tiny(**kw)
    s.flags.update(fl)
    orc = _oracle(s, nfft=64)
    for a in ("x1", "x2", "x3", "x1s", "x2s", "x3s", "d1", "d2", "d3", "d1s", "d2s", "d3s",
              "sigma11", "sigma22", "sigma33"):
        assert np.array_equal(orc.array(a), golden["%s_%s" % (name, a)]), a
    info = golden[name + "_info"]
    assert (orc.n1, orc.n2, orc.n3, orc.lE) == (int(info[0]), int(info[1]), int(info[2]), int(info[6]))
    assert orc.lextend == info[3] and orc.rho_water == info[4]
    orc.set_freq(0.75)
    x = _rand(orc.lE, 100)
    if name + "_apply" in golden:
        assert np.array_equal(orc.apply(x), golden[name + "_apply"])
    assert np.array_equal(orc.lusgs(x), golden[name + "_lusgs"])
    src = s.acq.sources[0]
    assert np.array_equal(orc.inject(0, *src[:3]), golden[name + "_inject"])
    rec = np.array([[orc.extract(ch, *r[:3], x) for r in s.acq.receivers] for ch in range(3)])
    assert np.array_equal(rec.astype(np.complex64), golden[name + "_extract"])


@pytest.mark.parametrize("preco,algopt", [(1, 1), (0, 1), (1, 2), (0, 2)])
def test_golden_solves(golden, preco, algopt):
    name, kw, fl = _cases()[0]
    s = surveys.tiny(**kw)
    s.flags.update(fl)
    orc = _oracle(s)
    orc.set_freq(0.75)
    tol = float(np.float32(np.float32(1e-8) / np.float32(0.75)))
    b = orc.inject(0, *s.acq.sources[0][:3])
    if algopt == 1:
        x, it, _ = orc.fbicgstab(b, 60, tol, precond=preco)
    elif preco:
        x, it, _ = orc.fgmres(b, 6, tol, 10)
    else:
        x, it, _ = orc.gmres(b, 6, tol, 10)
    assert np.array_equal(x, golden["%s_solve_p%d_a%d" % (name, preco, algopt)])


@pytest.mark.parametrize("case", [c[0] for c in _cases()])
def test_golden_other_preconditioners(golden, case):
    """aphi_apply (fdfd.c:513-821) on every case, airbc 0 and 1, fvm 0 and
    1; gmres_apply (gmres_init_free.c:58-137, mp = 8) where the reference
    build can run the operator (airbc = 0)."""
    name, kw, fl = [c for c in _cases() if c[0] == case][0]
    s = surveys.tiny(**kw)
    s.flags.update(fl)
    orc = _oracle(s, nfft=64)
    orc.set_freq(0.75)
    x = _rand(orc.lE, 100)
    assert np.array_equal(orc.aphi(x), golden[name + "_aphi"])
    if name + "_gmresprec" in golden:
        assert np.array_equal(orc.gmres_prec(x, mp=8), golden[name + "_gmresprec"])


@pytest.mark.parametrize("preco,algopt", [(2, 1), (2, 2), (3, 1), (3, 2)])
def test_golden_solves_other_preconditioners(golden, preco, algopt):
    """main.c:110-117: F-BiCGStab and FGMRES(10) with the inner-GMRES (mp=8)
    and A-phi preconditioners, 60 iterations, bit for bit"""
    name, kw, fl = _cases()[0]
    s = surveys.tiny(**kw)
    s.flags.update(fl)
    orc = _oracle(s)
    orc.set_freq(0.75)
    orc.set_mp(8)
    tol = float(np.float32(np.float32(1e-8) / np.float32(0.75)))
    b = orc.inject(0, *s.acq.sources[0][:3])
    if algopt == 1:
        x, it, _ = orc.fbicgstab(b, 60, tol, precond=preco)
    else:
        x, it, _ = orc.fgmres(b, 6, tol, 10, precond=preco)
    assert np.array_equal(x, golden["%s_solve_p%d_a%d" % (name, preco, algopt)])


@pytest.mark.skipif(not os.path.exists(O.REF_HARNESS), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("seed,airbc,fvm", [(11, 0, 0), (12, 1, 0), (13, 1, 1), (14, 0, 1)])
def test_live_reference(workdir, seed, airbc, fvm):
    s = surveys.tiny(seed, nx=7, ny=8, nz=6, airbc=airbc)
    s.flags.update(fvm=fvm, nb=3)
    surveys.write_files(workdir, s)
    argv = s.argv(workdir)
    orc = _oracle(s, nfft=64)
    O.run_ref(workdir, argv, "grid")
    for a in ("x1", "x3s", "d2", "sigma11", "sigma22", "sigma33"):
        assert np.array_equal(np.fromfile(os.path.join(workdir, "ref_" + a)), orc.array(a)), a
    f = 1.25
    orc.set_freq(f)
    x = _rand(orc.lE, seed)
    x.tofile(os.path.join(workdir, "xin"))
    if not airbc:
        O.run_ref(workdir, argv, "apply", hfreq=f, hin=os.path.join(workdir, "xin"))
        assert np.array_equal(np.fromfile(os.path.join(workdir, "ref_y"), np.complex128), orc.apply(x))
    O.run_ref(workdir, argv, "lusgs", hfreq=f, hin=os.path.join(workdir, "xin"))
    assert np.array_equal(np.fromfile(os.path.join(workdir, "ref_y"), np.complex128), orc.lusgs(x))


# ---------------------------------------------------- reference-held data

@pytest.fixture(scope="module")
def layered():
    return np.load(os.path.join(GOLDEN, "run_layered.npz"))


def test_run_layered_extended_model(layered):
    """extend_model on the run_layered inputs reproduces the reference's
    own param_extended=1 dumps (run.sh: nb=10, default lextend)."""
    g = layered
    nx, ny, nz = len(g["fx1"]) - 1, len(g["fx2"]) - 1, len(g["fx3"]) - 1
    rho = [np.broadcast_to(g[k + "_profile"][:, None, None], (nz, ny, nx)).astype(np.float32)
           for k in ("frho11", "frho22", "frho33")]
    orc = O.Oracle(g["fx1"], g["fx2"], g["fx3"], *rho, nb=10, airbc=1, nfft=64)
    for ax in (1, 2, 3):
        assert np.array_equal(orc.array("x%d" % ax), g["fx%d_extended" % ax])
    n1, n2, n3 = orc.n1, orc.n2, orc.n3f
    for c in (11, 22, 33):
        cell = orc.array("cell%d" % c).reshape(n3, n2, n1)
        assert np.all(cell == cell[:, :1, :1])
        assert np.array_equal(cell[:, 0, 0], g["frho%d_extended_profile" % c])


# ------------------------------------------------------------ mathematics

def test_fft2_matches_numpy():
    s = surveys.tiny(1, airbc=1)
    for n in (16, 32, 64):
        orc = _oracle(s, nfft=n)
        d = _rand(n * n, n).reshape(n, n)
        fwd = orc.fft2(d.ravel(), -1).reshape(n, n)
        # data[i1 + n*i2]: a 2-D DFT over both axes, e^{-i} forward
        assert np.allclose(fwd, np.fft.fft2(d), rtol=0, atol=1e-12 * np.abs(fwd).max())
        inv = orc.fft2(d.ravel(), +1).reshape(n, n)
        assert np.allclose(inv, np.fft.ifft2(d) * n * n, rtol=0, atol=1e-12 * np.abs(inv).max())


def test_airwave_multipliers_closed_form():
    s = surveys.tiny(4, airbc=1)
    N = 64
    orc = _oracle(s, nfft=N)
    d1fft = (orc.array("x1")[-1] - orc.array("x1")[0]) / (N - 1)
    d2fft = (orc.array("x2")[-1] - orc.array("x2")[0]) / (N - 1)
    dz = float(s.model.x3node[1] - s.model.x3node[0])
    kx = 2 * np.pi * np.fft.fftfreq(N, d1fft)
    ky = 2 * np.pi * np.fft.fftfreq(N, d2fft)
    # fftfreq puts N/2 at -N/2; the reference keeps i <= N/2 positive
    kx[N // 2] = -kx[N // 2]
    ky[N // 2] = -ky[N // 2]
    KX, KY = np.meshgrid(kx, ky)
    kz = np.hypot(KX, KY)
    d1s0 = orc.array("d1s")[0]
    d2s0 = orc.array("d2s")[0]
    s1 = np.exp(-kz * dz / 2) * 1j * KX / (kz + 1e-15) * np.exp(-1j * KX * d1s0 / 2)
    s2 = np.exp(-kz * dz / 2) * 1j * KY / (kz + 1e-15) * np.exp(-1j * KY * d2s0 / 2)
    assert np.allclose(orc.array("sH1").reshape(N, N), s1, rtol=0, atol=1e-13)
    assert np.allclose(orc.array("sH2").reshape(N, N), s2, rtol=0, atol=1e-13)


def _dense(orc, freq):
    orc.set_freq(freq)
    n = orc.lE
    A = np.zeros((n, n), complex)
    for q in range(n):
        e = np.zeros(n, complex)
        e[q] = 1
        A[:, q] = orc.apply(e)
    return A


def _interior(orc):
    """unknowns the LU-SGS sweeps update (fdfd.c:347-424 loop ranges)"""
    n1, n2, n3 = orc.n1, orc.n2, orc.n3
    mask = []
    for k in range(n3 + 1):
        for j in range(n2 + 1):
            for i in range(n1):
                mask.append(1 <= k <= n3 - 1 and 1 <= j <= n2 - 1)
    for k in range(n3 + 1):
        for j in range(n2):
            for i in range(n1 + 1):
                mask.append(1 <= k <= n3 - 1 and 1 <= i <= n1 - 1)
    for k in range(n3):
        for j in range(n2 + 1):
            for i in range(n1 + 1):
                mask.append(1 <= j <= n2 - 1 and 1 <= i <= n1 - 1)
    return np.array(mask)


def _small(airbc=0, fvm=0, seed=3):
    s = surveys.tiny(seed, nx=3, ny=3, nz=3, airbc=airbc)
    s.flags.update(nb=1, fvm=fvm, rho_air=2.0)
    return s


def _cell_index(orc):
    """(k, j, i) of every unknown in the [Ex | Ey | Ez] layout"""
    n1, n2, n3 = orc.n1, orc.n2, orc.n3
    out = []
    for shape in ((n3 + 1, n2 + 1, n1), (n3 + 1, n2, n1 + 1), (n3, n2 + 1, n1 + 1)):
        k, j, i = np.meshgrid(*[np.arange(d) for d in shape], indexing="ij")
        out.append(np.stack([k.ravel(), j.ravel(), i.ravel()], 1))
    return np.concatenate(out)


@pytest.mark.parametrize("fvm", [0, 1])
def test_lusgs_is_symmetric_gauss_seidel(fvm):
    """y = (D+U)^-1 D (D+L)^-1 x over the swept unknowns, ordered as the
    vector ([Ex lex; Ey lex; Ez lex]).  With fvm=1 every row -- Ex, Ey and
    Ez alike -- is scaled by d1s[i]*d2[j]*d3[k]*I*omega*mu0
    (fdfd.c:365,391,417), not by the component's own edge volume that
    fdfd_apply uses, so the dense matrix is built that way."""
    s = _small(fvm=0)
    orc0 = _oracle(s)
    A = _dense(orc0, 0.5)
    if fvm:
        s.flags["fvm"] = 1
        orc = _oracle(s)
        orc.set_freq(0.5)
        kji = _cell_index(orc)
        d1s, d2, d3 = orc.array("d1s"), orc.array("d2"), orc.array("d3")
        vol = d1s[kji[:, 2]] * d2[kji[:, 1]] * d3[kji[:, 0]]
        iomu = 1j * 2 * np.pi * float(np.float32(0.5)) * MU0
        A = (vol * iomu)[:, None] * A
    else:
        orc = orc0
    I = _interior(orc)
    AI = A[np.ix_(I, I)]
    D = np.diag(np.diag(AI))
    L, U = np.tril(AI, -1), np.triu(AI, 1)
    x = _rand(orc.lE, 3)
    y = orc.lusgs(x)
    yd = np.linalg.solve(D + U, D @ np.linalg.solve(D + L, x[I]))
    assert np.allclose(y[I], yd, rtol=0, atol=1e-10 * np.abs(yd).max())
    assert np.all(y[~I] == 0)


def test_curl_of_gradient_vanishes():
    s = surveys.tiny(8, nx=6, ny=6, nz=5, airbc=0)
    s.flags.update(nb=2)
    orc = _oracle(s)
    orc.set_freq(0.25)
    n1, n2, n3 = orc.n1, orc.n2, orc.n3
    rng = np.random.default_rng(0)
    phi = np.zeros((n3 + 1, n2 + 1, n1 + 1))
    phi[1:-1, 1:-1, 1:-1] = rng.standard_normal((n3 - 1, n2 - 1, n1 - 1))
    d1s, d2s, d3s = orc.array("d1s"), orc.array("d2s"), orc.array("d3s")
    ex = (phi[:, :, 1:] - phi[:, :, :-1]) / d1s[None, None, :n1]
    ey = (phi[:, 1:, :] - phi[:, :-1, :]) / d2s[None, :n2, None]
    ez = (phi[1:, :, :] - phi[:-1, :, :]) / d3s[:n3, None, None]
    x = np.concatenate([ex.ravel(), ey.ravel(), ez.ravel()]).astype(complex)
    y = orc.apply(x)
    I = _interior(orc)
    sig = np.concatenate([orc.array("sigma11"), orc.array("sigma22"), orc.array("sigma33")])
    want = -sig * x
    assert np.max(np.abs(y[I] - want[I])) <= 1e-9 * np.max(np.abs(want[I]))


def test_volume_scaled_operator_is_complex_symmetric():
    s = _small(fvm=1, seed=9)
    orc = _oracle(s)
    A = _dense(orc, 0.75)
    I = _interior(orc)
    AI = A[np.ix_(I, I)]
    assert np.max(np.abs(AI - AI.T)) <= 1e-12 * np.max(np.abs(AI))


@pytest.mark.parametrize("solver", ["fbicgstab", "fgmres"])
def test_converged_solve_matches_direct_solve(solver):
    s = _small(seed=5)
    orc = _oracle(s)
    A = _dense(orc, 1.25)
    I = _interior(orc)
    b = np.zeros(orc.lE, complex)
    b[I] = _rand(int(I.sum()), 1)
    if solver == "fbicgstab":
        x, it, hist = orc.fbicgstab(b, 2000, 1e-12)
    else:
        x, it, hist = orc.fgmres(b, 200, 1e-12, 10)
    xd = np.linalg.solve(A[np.ix_(I, I)], b[I])
    assert it < (2000 if solver == "fbicgstab" else 2000)
    assert np.max(np.abs(x[I] - xd)) <= 1e-8 * np.max(np.abs(xd))
    r = b - A @ x
    assert np.linalg.norm(r[I]) <= 1e-10 * np.linalg.norm(b)


def test_chunked_dot_is_a_reordering():
    s = surveys.tiny(1, airbc=0)
    orc = _oracle(s)
    a, b = _rand(50001, 1), _rand(50001, 2)
    exact = np.vdot(b, a)  # sum a * conj(b)
    orc.set_red(O.RED_SERIAL)
    ds = orc.dot(a, b)
    orc.set_red(O.RED_CHUNKED)
    dc = orc.dot(a, b)
    assert abs(ds - exact) <= 1e-12 * abs(exact) and abs(dc - exact) <= 1e-12 * abs(exact)
    # single-chunk tree on 256 values equals pairwise summation by hand
    v = _rand(256, 3)
    w = np.ones(256, complex)
    t = list(v)
    h = 128
    while h >= 1:
        t = [t[i] + t[i + h] for i in range(h)]
        h //= 2
    assert orc.dot(v, w) == t[0]
