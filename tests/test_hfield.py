"""H channels at the receivers (Hx, Hy, Hz).

The reference has no H channels: its extract_emf only knows Ex, Ey, Ez
(inject_extract.c:123-213) and the H entries of dcal_fd stay zero.  Here H
is formed from the solved E as fdfd_apply forms it (H = curl E / (i w mu0)
on the faces, fdfd.c:125-149) and interpolated trilinearly like the E
channels.  With no reference output to compare against, the host code and
the oracle are pinned by what the mathematics fixes:
  - the curl of a discrete gradient is zero (to roundoff), on any
    nonuniform grid and under the airwave boundary's plane offset;
  - a linear field gives a constant H equal to its slope / (i w mu0), with
    the sign and the derivative each component takes (Ey = a x -> Hz,
    Ez = b y -> Hx, Ex = c z -> Hy), and zero in the other two components;
  - mirror symmetry: for an Ex source on the plane y = 0 of a model
    symmetric in y, Hx and Hz vanish on that plane, Hy does not.
The host product equals the oracle bit for bit on random fields."""
import numpy as np
import pytest

from demo_matrixfree_csem_amd import surveys
from demo_matrixfree_csem_amd.session import Session
import oracle as O

CH6 = "Ex,Ey,Ez,Hx,Hy,Hz"


def _pair(workdir, airbc=1, seed=3, survey=None):
    s = survey or surveys.tiny(seed, airbc=airbc)
    s.flags.update(n1fft=64, n2fft=64, chrec=CH6)
    surveys.write_files(workdir, s)
    ses = Session(s.argv(workdir), device=False)
    m = s.model
    orc = O.Oracle(m.x1node, m.x2node, m.x3node, m.rho11, m.rho22, m.rho33, nb=s.flags["nb"],
                   airbc=s.flags["airbc"], fvm=s.flags.get("fvm", 0), nfft=64,
                   rho_air=float(s.flags.get("rho_air", 1e6)))
    return s, ses, orc


def _orc_channels(orc, s, x):
    return np.array([[orc.extract(ch, *r[:3], x) for r in s.acq.receivers] for ch in range(6)])


@pytest.mark.parametrize("airbc", [0, 1])
def test_h_channels_host_equals_oracle(workdir, airbc):
    s, ses, orc = _pair(workdir, airbc)
    rng = np.random.default_rng(11)
    for f in s.freqs:
        ses.set_freq(f)
        orc.set_freq(f)
        x = rng.standard_normal(ses.n) + 1j * rng.standard_normal(ses.n)
        got = ses.extract(x)
        want = _orc_channels(orc, s, x)
        assert got.shape == (6, len(s.acq.receivers))
        assert np.all(np.isfinite(want)) and np.all(np.abs(want[3:]) > 0)
        assert np.array_equal(got, want.astype(np.complex64))


def _fields(orc, airbc, nb):
    """Operator-grid shapes and the spacings the operator uses:
    d*s[i] = x[i+1] - x[i] (extend_model.c:8-21), vertical ones of operator
    plane k at full-grid index k + nb under the airwave boundary."""
    n1, n2, n3 = orc.n1, orc.n2, orc.n3
    ks = nb if airbc else 0
    d1s, d2s = orc.array("d1s")[:n1], orc.array("d2s")[:n2]
    d3s = orc.array("d3s")[ks:ks + n3]
    x1, x2, x3 = orc.array("x1")[:n1 + 1], orc.array("x2")[:n2 + 1], orc.array("x3")[ks:ks + n3 + 1]
    return n1, n2, n3, d1s, d2s, d3s, x1, x2, x3


def _pack(Ex, Ey, Ez):
    return np.concatenate([Ex.ravel(), Ey.ravel(), Ez.ravel()])


@pytest.mark.parametrize("airbc", [0, 1])
def test_h_of_a_gradient_is_zero(workdir, airbc):
    s, ses, orc = _pair(workdir, airbc)
    f = 0.75
    ses.set_freq(f)
    orc.set_freq(f)
    n1, n2, n3, d1s, d2s, d3s, *_ = _fields(orc, airbc, s.flags["nb"])
    rng = np.random.default_rng(5)
    phi = rng.standard_normal((n3 + 1, n2 + 1, n1 + 1)) + 1j * rng.standard_normal((n3 + 1, n2 + 1, n1 + 1))
    Ex = (phi[:, :, 1:] - phi[:, :, :-1]) / d1s[None, None, :]
    Ey = (phi[:, 1:, :] - phi[:, :-1, :]) / d2s[None, :, None]
    Ez = (phi[1:, :, :] - phi[:-1, :, :]) / d3s[:, None, None]
    x = _pack(Ex, Ey, Ez)
    assert x.size == ses.n
    # scale of one difference term of H: |E| / spacing / (w mu0)
    w = 2 * np.pi * float(np.float32(f)) * 4e-7 * np.pi
    scale = np.abs(x).max() / min(d1s.min(), d2s.min(), d3s.min()) / w
    got = ses.extract(x)
    want = _orc_channels(orc, s, x)
    assert np.array_equal(got, want.astype(np.complex64))
    assert np.abs(want[3:]).max() < 1e-12 * scale
    # a random field of the same size is not curl-free
    xr = rng.standard_normal(ses.n) * np.abs(x).max()
    assert np.abs(_orc_channels(orc, s, xr)[3:]).max() > 1e-3 * scale


@pytest.mark.parametrize("airbc", [0, 1])
def test_h_of_linear_fields(workdir, airbc):
    s, ses, orc = _pair(workdir, airbc)
    f = 1.25
    ses.set_freq(f)
    orc.set_freq(f)
    n1, n2, n3, d1s, d2s, d3s, x1, x2, x3 = _fields(orc, airbc, s.flags["nb"])
    iwmu = 1.0 / (1j * 2 * np.pi * float(np.float32(f)) * 4e-7 * np.pi)
    zx = np.zeros((n3 + 1, n2 + 1, n1))
    zy = np.zeros((n3 + 1, n2, n1 + 1))
    zz = np.zeros((n3, n2 + 1, n1 + 1))
    a, b, c = 0.7 - 0.2j, -1.3 + 0.4j, 0.25 + 0.9j
    cases = [  # (E, channel expected nonzero, value)
        (_pack(zx, a * np.broadcast_to(x1[None, None, :], zy.shape), zz), 5, a * iwmu),  # Ey = a x -> Hz
        (_pack(zx, zy, b * np.broadcast_to(x2[None, :, None], zz.shape)), 3, b * iwmu),  # Ez = b y -> Hx
        (_pack(c * np.broadcast_to(x3[:, None, None], zx.shape), zy, zz), 4, c * iwmu),  # Ex = c z -> Hy
    ]
    for x, ch, v in cases:
        want = _orc_channels(orc, s, x)
        assert np.array_equal(ses.extract(x), want.astype(np.complex64))
        assert np.abs(want[ch] - v).max() < 1e-9 * abs(v), (ch, want[ch], v)
        for o in {3, 4, 5} - {ch}:
            assert np.abs(want[o]).max() < 1e-9 * abs(v), (o, want[o])


def _symmetric_survey():
    """Uniform 100 m cells, x and y in [-500, 500] (y = 0 a node), layered
    below a water layer with a resistor that is mirror-symmetric in y; an
    Ex source and receivers on y = 0.  No airwave boundary: its transform
    grid and half-cell phase factors (airbc_init_free.c:69-105) are not
    mirror-symmetric, so the operator is not either; instead the padding
    above the water is given a conductive value (rho_air = 1), which keeps
    the solve well conditioned."""
    nx, ny, nz = 10, 10, 7
    x = (-500.0 + 100.0 * np.arange(nx + 1)).astype(np.float32)
    y = (-500.0 + 100.0 * np.arange(ny + 1)).astype(np.float32)
    z = (100.0 * np.arange(nz + 1)).astype(np.float32)
    rng = np.random.default_rng(8)
    r = np.broadcast_to(10 ** rng.uniform(0, 1.5, (nz, 1, 1)), (nz, ny, nx)).copy()
    r[4, 3:7, 2:6] = 50.0  # cells 3..6 of 0..9 along y: symmetric about y = 0
    zc = 0.5 * (z[1:] + z[:-1])
    r = np.where((zc < 200)[:, None, None], surveys.RHO_WATER, r).astype(np.float32)
    src = [(-180.0, 0.0, 145.0)]
    rec = [(xr, 0.0, 200.0) for xr in (-350.0, -60.0, 130.0, 320.0)]
    return surveys.Survey("sym", surveys.Model(x, y, z, r, r, r),
                          surveys.Acquisition(surveys._acq_table(src), surveys._acq_table(rec)), [0.75],
                          dict(fvm=0, airbc=0, rho_air=1.0, preco=1, algopt=1, tol=1e-10, niter=1500, nb=3,
                               chsrc="Ex", chrec=CH6))


def test_mirror_symmetry_of_a_solved_field(workdir):
    s, ses, orc = _pair(workdir, survey=_symmetric_survey())
    f = s.freqs[0]
    orc.set_freq(f)
    ses.set_freq(f)
    b = orc.inject(0, *s.acq.sources[0][:3])
    x, its = orc.fbicgstab(b, 1500, 1e-10, precond=1)[:2]
    assert 0 < its < 1500
    h = _orc_channels(orc, s, x)
    assert np.array_equal(ses.extract(x), h.astype(np.complex64))
    hy = np.abs(h[4]).min()
    assert hy > 0
    # Hx and Hz are odd in y, Hy is even
    assert np.abs(h[3]).max() < 1e-9 * hy and np.abs(h[5]).max() < 1e-9 * hy


def test_h_outside_the_grid_is_nan(workdir, capfd):
    """A receiver whose eight H faces are not all inside the operator grid
    gets NaN for that channel and a message.  Here it is 10 m below the sea
    surface, where the operator grid starts under the airwave boundary: Hx
    and Hy (at cell centres in z) would need a plane above the grid, Hz (on
    the z nodes, the surface plane included) does not; the E channels are
    unaffected."""
    s, ses, orc = _pair(workdir, airbc=1)
    ses.set_freq(0.5)
    orc.set_freq(0.5)
    x = np.random.default_rng(2).standard_normal(ses.n) + 0j
    r0 = s.acq.receivers[0]
    top = (float(r0[0]), float(r0[1]), 10.0)  # 10 m below the surface
    got = ses.extract_at(x, [top, tuple(float(v) for v in r0[:3])])
    assert np.all(np.isnan(got[3:5, 0])) and np.isfinite(got[5, 0]) and np.all(np.isfinite(got[3:, 1]))
    assert np.all(np.isfinite(got[:3, 0]))
    assert np.isnan(orc.extract(4, *top, x))
    assert "too close to the grid's top or bottom" in capfd.readouterr().err
