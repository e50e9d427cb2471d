"""Synthetic marine CSEM surveys in the reference's input formats.

A survey is a model (float32 node coordinates and float32 cell
resistivities, x fastest, as emf_init reads them: emf_init_free.c:77-150)
plus an acquisition (text source/receiver tables, acq_init_free.c:38-80).
`write_files` writes exactly those files so that the reference's own
command line (`fx1= ... frho11= ... fsrc= frec=`) and this package's
driver both read them.

Generators cover the measured workloads:
  layered_64      64x64x64 uniform 100 m cells, water + 1 ohm-m + 100 ohm-m slab
  layered_100     100^3 uniform 100 m cells, 8 seeded log-uniform layers
  marine          any grid, layered or Gaussian-random-field resistivity under water
  vti_marine      anisotropic random-field model with resistive ellipsoids, six channels
  topo_marine     seafloor topography over layering, eight sources along a line
  serve_request   a seeded serving request: model, frequencies, sources, receivers
  tiny            small seeded random models for parity tests
"""
from dataclasses import dataclass, field
import os

import numpy as np

RHO_WATER = 0.3125


@dataclass
class Model:
    x1node: np.ndarray
    x2node: np.ndarray
    x3node: np.ndarray
    rho11: np.ndarray  # (nz, ny, nx) float32
    rho22: np.ndarray
    rho33: np.ndarray

    @property
    def shape(self):
        return len(self.x1node) - 1, len(self.x2node) - 1, len(self.x3node) - 1


@dataclass
class Acquisition:
    sources: np.ndarray  # (ns, 5) float32: x, y, z, azimuth, dip
    receivers: np.ndarray  # (nr, 5)


@dataclass
class Survey:
    name: str
    model: Model
    acq: Acquisition
    freqs: list
    flags: dict = field(default_factory=dict)

    def argv(self, workdir=".", isrc=0):
        """key=value flags for the reference command line (README.md:49-112)."""
        f = dict(self.flags)
        f["freqs"] = ",".join("%g" % v for v in self.freqs)
        for k in ("fx1", "fx2", "fx3", "frho11", "frho22", "frho33"):
            f[k] = os.path.join(workdir, k)
        f["fsrc"] = os.path.join(workdir, "sources%d.txt" % isrc if isrc else "sources.txt")
        f["frec"] = os.path.join(workdir, "receivers.txt")
        return ["%s=%s" % (k, v) for k, v in f.items()]


def _acq_table(rows):
    a = np.asarray(rows, np.float32)
    if a.ndim == 1:
        a = a[None, :]
    if a.shape[1] == 3:
        a = np.concatenate([a, np.zeros((a.shape[0], 2), np.float32)], 1)
    return a


def write_table(path, rows, label):
    with open(path, "w") as f:
        f.write("          x1          x2          x3     azimuth         dip         %s\n" % label)
        for i, r in enumerate(rows):
            # %.9g round-trips float32 through the reference's fscanf("%f")
            f.write("%16.9g %16.9g %16.9g %12.9g %12.9g %8d\n" % (tuple(float(v) for v in r[:5]) + (i + 1,)))


def write_files(workdir, survey):
    os.makedirs(workdir, exist_ok=True)
    m = survey.model
    for k, a in (("fx1", m.x1node), ("fx2", m.x2node), ("fx3", m.x3node)):
        np.asarray(a, np.float32).tofile(os.path.join(workdir, k))
    for k, a in (("frho11", m.rho11), ("frho22", m.rho22), ("frho33", m.rho33)):
        np.ascontiguousarray(a, np.float32).tofile(os.path.join(workdir, k))
    write_table(os.path.join(workdir, "sources.txt"), survey.acq.sources, "iTx")
    for i in range(1, len(survey.acq.sources)):
        write_table(os.path.join(workdir, "sources%d.txt" % i), survey.acq.sources[i:i + 1], "iTx")
    write_table(os.path.join(workdir, "receivers.txt"), survey.acq.receivers, "iRx")


def _centers(nodes):
    nodes = np.asarray(nodes, np.float64)
    return 0.5 * (nodes[1:] + nodes[:-1])


def layered_64():
    """64^3 uniform 100 m cells, seawater above 1000 m, 1 ohm-m below and a
    100 ohm-m slab at z 2000-2100 m, |x|,|y| <= 1500 m; one Ex source at
    (-2000, 0, 950), 33 receivers on y=0, z=1000; 0.25 Hz."""
    x = (-3200 + 100 * np.arange(65)).astype(np.float32)
    z = (100 * np.arange(65)).astype(np.float32)
    xc, zc = _centers(x), _centers(z)
    rho = np.where(zc[:, None, None] < 1000, RHO_WATER, 1.0) * np.ones((64, 64, 64))
    slab = (np.abs(xc)[None, None, :] <= 1500) & (np.abs(xc)[None, :, None] <= 1500) & \
           (zc[:, None, None] > 2000) & (zc[:, None, None] < 2100)
    rho = np.where(slab, 100.0, rho).astype(np.float32)
    rec = [(xr, 0.0, 1000.0) for xr in np.arange(-3200, 3201, 200)]
    return Survey("layered_64", Model(x, x.copy(), z, rho, rho.copy(), rho.copy()),
                  Acquisition(_acq_table([(-2000.0, 0.0, 950.0)]), _acq_table(rec)), [0.25],
                  dict(fvm=0, airbc=1, preco=1, algopt=1, tol=1e-8, niter=1000, nb=10,
                       lextend=40e3, chsrc="Ex", chrec="Ex,Ey,Ez"))


def layered_64_variant(nx, ny, nz, sources=(-2000.0, -1000.0, 1000.0, 2000.0), freqs=(0.25, 0.75)):
    """The 64^3 parity model's construction on another grid of 100 m cells:
    x and y centred on 0, z from 0; seawater above 1000 m, 1 ohm-m below,
    the 100 ohm-m slab at z 2000-2100 m for |x|, |y| <= 1500 m; Ex sources
    at (x, 0, 950) for x in `sources`; receivers every 200 m on y = 0,
    z = 1000 within the grid."""
    x = (100.0 * (np.arange(nx + 1) - nx / 2)).astype(np.float32)
    y = (100.0 * (np.arange(ny + 1) - ny / 2)).astype(np.float32)
    z = (100.0 * np.arange(nz + 1)).astype(np.float32)
    xc, yc, zc = _centers(x), _centers(y), _centers(z)
    rho = np.where(zc[:, None, None] < 1000, RHO_WATER, 1.0) * np.ones((nz, ny, nx))
    slab = (np.abs(xc)[None, None, :] <= 1500) & (np.abs(yc)[None, :, None] <= 1500) & \
           (zc[:, None, None] > 2000) & (zc[:, None, None] < 2100)
    rho = np.where(slab, 100.0, rho).astype(np.float32)
    half = 100.0 * (nx // 2)
    rec = [(xr, 0.0, 1000.0) for xr in np.arange(-half, half + 1, 200)]
    return Survey("layered_%dx%dx%d" % (nx, ny, nz), Model(x, y, z, rho, rho.copy(), rho.copy()),
                  Acquisition(_acq_table([(sx, 0.0, 950.0) for sx in sources]), _acq_table(rec)), list(freqs),
                  dict(fvm=0, airbc=1, preco=1, algopt=1, tol=1e-8, niter=2000, nb=10,
                       lextend=40e3, chsrc="Ex", chrec="Ex,Ey,Ez"))


def layered_100(seed=1):
    """100^3 uniform 100 m cells on x,y in [-5000,5000], z in [0,10000]:
    water (0.3125 ohm-m) above 1000 m, then 8 layers of 1125 m with seeded
    log-uniform resistivity in [0.5, 200] ohm-m (by cell centre).  One Ex
    source at (0,0,950), 51 receivers on y=0, z=1000; 0.25/0.75/1.25 Hz;
    all six channels (the H channels are beyond the reference)."""
    rng = np.random.default_rng(seed)
    x = (-5000 + 100 * np.arange(101)).astype(np.float32)
    z = (100 * np.arange(101)).astype(np.float32)
    zc = _centers(z)
    rl = 10 ** rng.uniform(np.log10(0.5), np.log10(200.0), 8)
    layer = np.clip(((zc - 1000) // 1125).astype(int), 0, 7)
    prof = np.where(zc < 1000, RHO_WATER, rl[layer])
    rho = (prof[:, None, None] * np.ones((100, 100, 100))).astype(np.float32)
    rec = [(xr, 0.0, 1000.0) for xr in np.arange(-5000, 5001, 200)]
    return Survey("layered_100", Model(x, x.copy(), z, rho, rho.copy(), rho.copy()),
                  Acquisition(_acq_table([(0.0, 0.0, 950.0)]), _acq_table(rec)), [0.25, 0.75, 1.25],
                  dict(fvm=0, airbc=1, preco=1, algopt=1, tol=1e-8, niter=1000, nb=10,
                       chsrc="Ex", chrec="Ex,Ey,Ez,Hx,Hy,Hz"))


def tiny(seed=0, nx=10, ny=9, nz=7, airbc=1, aniso=True, dz=100.0):
    """A small seeded model with random cell sizes and resistivities, a
    water layer on top (so rho_water is the water value), one Ex source
    and a few receivers.  nz <= ny keeps the operator's H scratch layout
    unaliased (the reference sizes it with n2 for the k extent)."""
    rng = np.random.default_rng(seed)

    def axis(n, d0, origin):
        d = rng.uniform(0.6, 1.4, n) * d0
        return (origin + np.concatenate([[0.0], np.cumsum(d)])).astype(np.float32)

    x = axis(nx, 100.0, -50.0 * nx)
    y = axis(ny, 100.0, -50.0 * ny)
    z = (dz * np.arange(nz + 1)).astype(np.float32)
    zc = _centers(z)
    r11 = 10 ** rng.uniform(0, 1.5, (nz, ny, nx))
    r22 = r11 if not aniso else 10 ** rng.uniform(0, 1.5, (nz, ny, nx))
    r33 = 2 * r11 if aniso else r11
    water = (zc < 2 * dz)[:, None, None]
    r11 = np.where(water, RHO_WATER, r11).astype(np.float32)
    r22 = np.where(water, RHO_WATER, r22).astype(np.float32)
    r33 = np.where(water, RHO_WATER, r33).astype(np.float32)
    xm, ym = float(x[0]), float(y[0])
    xw, yw = float(x[-1] - x[0]), float(y[-1] - y[0])
    src = [(xm + 0.31 * xw, ym + 0.47 * yw, 1.45 * dz)]
    rec = [(xm + f * xw, ym + 0.52 * yw, 2.0 * dz) for f in (0.2, 0.45, 0.6, 0.85)]
    return Survey("tiny%d" % seed, Model(x, y, z, r11, r22, r33),
                  Acquisition(_acq_table(src), _acq_table(rec)), [0.5, 1.25],
                  dict(fvm=0, airbc=airbc, preco=1, algopt=1, tol=1e-8, niter=200, nb=3,
                       chsrc="Ex", chrec="Ex,Ey,Ez"))


def vti_marine(seed=2, nx=180, ny=120, nz=120, nrec=121):
    """Anisotropic (VTI) marine model: per-axis cell sizes drawn uniformly
    in [50, 250] m (ascending nodes; x and y centred on 0, z from 0), sea
    water (0.3125 ohm-m) above 1000 m, below it rho11 = rho22 a log-normal
    Gaussian random field with median 2 ohm-m (FFT-filtered white noise,
    correlation length ~1 km, log-standard deviation 0.5) carrying three
    seeded resistive ellipsoids of 50-300 ohm-m, and rho33 = 2 rho11.
    One Ex source 50 m above the seafloor at the centre, `nrec` receivers
    on the seafloor along y = 0; nb=10, lextend=120 km; six frequencies
    0.125-1.25 Hz; all six channels (the H channels are beyond the
    reference).  The default size is 180 x 120 x 120 cells."""
    rng = np.random.default_rng(seed)

    def axis(n, centred):
        d = rng.uniform(50.0, 250.0, n)
        x = np.concatenate([[0.0], np.cumsum(d)])
        return (x - x[-1] / 2 if centred else x).astype(np.float32)

    x, y, z = axis(nx, True), axis(ny, True), axis(nz, False)
    xc, yc, zc = _centers(x), _centers(y), _centers(z)
    h = float(np.mean(np.diff(z)))
    w = rng.standard_normal((nz, ny, nx))
    k = [np.fft.fftfreq(nn, d=h) for nn in (nz, ny, nx)]
    kk = np.sqrt(k[0][:, None, None] ** 2 + k[1][None, :, None] ** 2 + k[2][None, None, :] ** 2)
    g = np.real(np.fft.ifftn(np.fft.fftn(w) * np.exp(-(kk * 1000.0) ** 2)))
    g = (g - g.mean()) / (g.std() + 1e-30)
    rho = 2.0 * np.exp(0.5 * g)
    below = zc > 1000.0
    zb = float(zc[below].min()), float(zc[-1])
    for _ in range(3):
        c = (rng.uniform(0.6 * x[0], 0.6 * x[-1]), rng.uniform(0.6 * y[0], 0.6 * y[-1]),
             rng.uniform(zb[0] + 0.2 * (zb[1] - zb[0]), zb[0] + 0.7 * (zb[1] - zb[0])))
        ax = (rng.uniform(0.1, 0.25) * (x[-1] - x[0]), rng.uniform(0.1, 0.25) * (y[-1] - y[0]),
              rng.uniform(0.03, 0.08) * (zb[1] - zb[0]))
        inside = (((zc[:, None, None] - c[2]) / ax[2]) ** 2 + ((yc[None, :, None] - c[1]) / ax[1]) ** 2 +
                  ((xc[None, None, :] - c[0]) / ax[0]) ** 2) <= 1.0
        rho = np.where(inside, rng.uniform(50.0, 300.0), rho)
    water = (~below)[:, None, None]
    r11 = np.where(water, RHO_WATER, rho).astype(np.float32)
    r33 = np.where(water, RHO_WATER, 2.0 * rho).astype(np.float32)
    half = 0.4 * float(x[-1] - x[0])
    rec = [(float(v), 0.0, 1000.0) for v in np.linspace(-half, half, nrec)]
    return Survey("vti_marine_%dx%dx%d" % (nx, ny, nz), Model(x, y, z, r11, r11.copy(), r33),
                  Acquisition(_acq_table([(0.0, 0.0, 950.0)]), _acq_table(rec)),
                  [0.125, 0.25, 0.5, 0.75, 1.0, 1.25],
                  dict(fvm=0, airbc=1, preco=1, algopt=1, nrestart=10, tol=1e-8, niter=2000, nb=10,
                       lextend=120e3, chsrc="Ex", chrec="Ex,Ey,Ez,Hx,Hy,Hz"))


def topo_marine(seed=3, nx=125, ny=125, nz=100, dx=100.0, nsrc=8, nrec=61):
    """Marine model with seafloor topography: uniform `dx` cells, x and y
    centred on 0, z from 0; the seafloor at 1000 m plus four seeded
    Gaussian bumps of up to +-300 m; sea water (0.3125 ohm-m) above it,
    below it seeded layers of 1-20 ohm-m by depth under the seafloor.
    `nsrc` Ex sources from x = -3500 to 3500 m on y = 0, 50 m above the
    local seafloor; `nrec` receivers on the seafloor along y = 0; six
    frequencies 0.125-1.25 Hz.  The default size is 125 x 125 x 100 cells."""
    rng = np.random.default_rng(seed)
    x = (dx * (np.arange(nx + 1) - nx / 2)).astype(np.float32)
    y = (dx * (np.arange(ny + 1) - ny / 2)).astype(np.float32)
    z = (dx * np.arange(nz + 1)).astype(np.float32)
    xc, yc, zc = _centers(x), _centers(y), _centers(z)
    bumps = [(rng.uniform(-0.35, 0.35) * (x[-1] - x[0]), rng.uniform(-0.35, 0.35) * (y[-1] - y[0]),
              rng.uniform(800.0, 2000.0), rng.choice([-1.0, 1.0]) * rng.uniform(100.0, 300.0)) for _ in range(4)]

    def seafloor(px, py):
        d = 1000.0 * np.ones(np.broadcast(px, py).shape)
        for bx, by, w, a in bumps:
            d = d + a * np.exp(-((px - bx) ** 2 + (py - by) ** 2) / (2 * w * w))
        return np.clip(d, 700.0, 1300.0)

    sf = seafloor(xc[None, :], yc[:, None])  # (ny, nx)
    nl = 6
    thick = rng.uniform(300.0, 1200.0, nl)
    rl = 10 ** rng.uniform(0.0, np.log10(20.0), nl)
    below = zc[:, None, None] - sf[None, :, :]  # depth under the seafloor
    layer = np.searchsorted(np.cumsum(thick), np.maximum(below, 0.0), side="right").clip(0, nl - 1)
    rho = np.where(below < 0, RHO_WATER, rl[layer]).astype(np.float32)
    sx = np.linspace(-3500.0, 3500.0, nsrc)
    src = [(float(v), 0.0, float(seafloor(v, 0.0)) - 50.0) for v in sx]
    half = 0.4 * float(x[-1] - x[0])
    rec = [(float(v), 0.0, float(seafloor(v, 0.0))) for v in np.linspace(-half, half, nrec)]
    return Survey("topo_marine_%dx%dx%d" % (nx, ny, nz), Model(x, y, z, rho, rho.copy(), rho.copy()),
                  Acquisition(_acq_table(src), _acq_table(rec)), [0.125, 0.25, 0.5, 0.75, 1.0, 1.25],
                  dict(fvm=0, airbc=1, preco=1, algopt=1, tol=1e-8, niter=2000, nb=10, chsrc="Ex",
                       chrec="Ex,Ey,Ez"))


GENERATORS = {"layered_64": layered_64, "layered_100": layered_100, "tiny": tiny, "vti_marine": vti_marine,
              "topo_marine": topo_marine}


def marine(seed, nx, ny, nz, kind="layered", median=2.0, dx=100.0, water=1000.0, nlayers=8):
    """Uniform dx cells centred in x and y, z from 0 down; seawater above
    `water` m, below it either `nlayers` seeded log-uniform layers around
    `median` (kind "layered") or a log-normal Gaussian random field with
    median `median` (FFT-filtered white noise, correlation length ~1 km,
    log-standard deviation 0.5; kind "grf")."""
    rng = np.random.default_rng(seed)
    x = (dx * (np.arange(nx + 1) - nx / 2)).astype(np.float32)
    y = (dx * (np.arange(ny + 1) - ny / 2)).astype(np.float32)
    z = (dx * np.arange(nz + 1)).astype(np.float32)
    zc = _centers(z)
    if kind == "grf":
        w = rng.standard_normal((nz, ny, nx))
        k = [np.fft.fftfreq(nn, d=dx) for nn in (nz, ny, nx)]
        kk = np.sqrt(k[0][:, None, None] ** 2 + k[1][None, :, None] ** 2 + k[2][None, None, :] ** 2)
        g = np.real(np.fft.ifftn(np.fft.fftn(w) * np.exp(-(kk * 1000.0) ** 2)))
        g = (g - g.mean()) / (g.std() + 1e-30)
        rho = median * np.exp(0.5 * g)
    else:
        depth = max(nz * dx - water, dx)
        rl = median * 10 ** rng.uniform(-1, 1, nlayers)
        layer = np.clip(((zc - water) // (depth / nlayers)).astype(int), 0, nlayers - 1)
        rho = rl[layer][:, None, None] * np.ones((nz, ny, nx))
    rho = np.where((zc < water)[:, None, None], RHO_WATER, rho).astype(np.float32)
    return Model(x, y, z, rho, rho.copy(), rho.copy())


# grids of the serving stream (cells): 64^3, 100^3, 128x128x96, 180x120x120
SERVE_GRIDS = [(64, 64, 64), (100, 100, 100), (128, 128, 96), (180, 120, 120)]


def serve_request(seed, grids=None, max_freqs=6, max_sources=4, nrec=(33, 121), n_models=4, nfft=None):
    """A seeded serving request: one of `n_models` seeded models per grid
    (layered or random-field, water median 1-5 ohm-m; requests share
    models, so their coefficients can stay on the device), 1..max_freqs
    frequencies in [0.125, 1.25] Hz, 1..max_sources Ex sources 50 m above
    the seafloor and 33..121 receivers on the seafloor along y = 0;
    nfft sets the airwave transform size (default the reference's 1024).
    Returns (model_key, model builder, freqs, sources, receivers)."""
    rng = np.random.default_rng(seed)
    grids = grids or SERVE_GRIDS
    g = grids[rng.integers(len(grids))]
    mseed = int(rng.integers(n_models))
    kind = "grf" if mseed % 2 else "layered"
    median = 1.0 + mseed  # 1..n_models ohm-m
    key = (g, kind, mseed)
    nf = int(rng.integers(1, max_freqs + 1))
    freqs = sorted(float(np.float32(v)) for v in rng.choice(
        [0.125, 0.25, 0.375, 0.5, 0.625, 0.75, 0.875, 1.0, 1.125, 1.25], nf, replace=False))
    half = 0.35 * 100.0 * g[0]
    ns = int(rng.integers(1, max_sources + 1))
    srcs = [(float(v), 0.0, 950.0) for v in np.linspace(-half / 2, half / 2, ns)]
    nr = int(rng.integers(nrec[0], nrec[1] + 1))
    recs = [(float(v), 0.0, 1000.0) for v in np.linspace(-half, half, nr)]

    flags = dict(fvm=0, airbc=1, preco=1, algopt=1, tol=1e-8, niter=2000, nb=10, chsrc="Ex", chrec="Ex,Ey,Ez")
    if nfft:
        flags.update(n1fft=nfft, n2fft=nfft)

    def build():
        return Survey("serve_%dx%dx%d_%s%d" % (g + (kind, mseed)),
                      marine(1000 + mseed, g[0], g[1], g[2], kind=kind, median=median),
                      Acquisition(_acq_table(srcs[:1]), _acq_table(recs)), freqs, dict(flags))
    return key, build, freqs, srcs, recs
