"""The topography workload (125 x 125 x 100 cells, 8 sources x 6
frequencies): its generator follows the stated construction, and (on the
GPU, at a reduced size) a batch of four (source, frequency) right-hand
sides solved in one mixed solve matches the FP64 drop-in solve of each."""
import numpy as np
import pytest

from demo_matrixfree_csem_amd import surveys


def test_topo_marine_construction():
    s = surveys.topo_marine()
    m = s.model
    assert m.shape == (125, 125, 100)
    zc = 0.5 * (m.x3node[1:] + m.x3node[:-1]).astype(np.float64)
    water = m.rho11 == surveys.RHO_WATER
    # water is the top of every column, down to the seafloor
    nw = water.sum(axis=0)
    assert np.array_equal(water, np.arange(len(zc))[:, None, None] < nw[None])
    depth = zc[nw - 1] + 50.0  # the seafloor lies within the last water cell's lower half
    assert depth.min() >= 650.0 and depth.max() <= 1350.0 and depth.max() - depth.min() > 200.0
    below = m.rho11[~water]
    assert below.min() >= 1.0 and below.max() <= 20.0 and len(np.unique(below)) >= 3
    assert np.array_equal(m.rho11, m.rho22) and np.array_equal(m.rho11, m.rho33)
    src = s.acq.sources
    assert len(src) == 8 and src[0][0] == -3500.0 and src[-1][0] == 3500.0 and np.all(src[:, 1] == 0.0)
    for sx, sy, sz in src[:, :3]:
        i = np.searchsorted(m.x1node, sx) - 1
        j = np.searchsorted(m.x2node, sy) - 1
        k = np.searchsorted(m.x3node, sz) - 1
        assert m.rho11[k, j, i] == surveys.RHO_WATER  # 50 m above the local seafloor: in the water
    assert s.freqs == [0.125, 0.25, 0.5, 0.75, 1.0, 1.25]


def _tol(f):
    return float(np.float32(np.float32(1e-8) / np.float32(f)))


def _rec_err(a, b):
    scale = np.abs(b).max(axis=1, keepdims=True)
    return float((np.abs(a - b) / np.where(scale > 0, scale, 1)).max())


@pytest.mark.gpu
def test_topo_batch_matches_fp64(workdir):
    """48 x 48 x 40 cells of 200 m: sources 2 and 5 (x = -1500 and 1500 m)
    at 0.5 and 1.25 Hz as one batch of four right-hand sides."""
    from demo_matrixfree_csem_amd import _lib
    from demo_matrixfree_csem_amd.session import Session
    s = surveys.topo_marine(3, 48, 48, 40, dx=200.0, nrec=33)
    surveys.write_files(workdir, s)
    print("case topo_marine 48x48x40 batch of 4 vs fp64", flush=True)
    ses = Session(s.argv(workdir))
    try:
        pairs = [(2, 0.5), (5, 0.5), (2, 1.25), (5, 1.25)]
        bs, fs, tols = [], [], []
        for isrc, f in pairs:
            ses.set_freq(f)
            bs.append(ses.inject_at(*s.acq.sources[isrc][:3]))
            fs.append(f)
            tols.append(_tol(f))
        xs, its, rel, nconv = ses.solve_mixed(bs, fs, tols, _lib.MixedOpts.default(max_inner=4000))
        assert nconv == 4, (its, rel)
        for r, (isrc, f) in enumerate(pairs):
            ses.set_freq(f)
            xp, itp = ses.solve(bs[r], tols[r], niter=4000, algopt=1)
            assert 0 < itp < 4000
            res = np.linalg.norm(bs[r] - ses.fdfd_apply(xs[r])) / np.linalg.norm(bs[r])
            assert res < tols[r], (f, res)
            a, b = ses.extract(xs[r]).astype(np.complex128), ses.extract(xp).astype(np.complex128)
            per_ch = np.abs(a - b).max(axis=1) / np.abs(b).max(axis=1)
            overall = np.abs(a - b).max() / np.abs(b).max()
            print("source %d f=%g: mixed %d, fp64 %d iterations; receiver error per channel %s, of the largest "
                  "field %.2e" % (isrc, f, its[r], itp, np.array2string(per_ch, precision=2), overall), flush=True)
            # Ex and Ez against their own maxima; Ey, small on the source's
            # line (the bumps only break the mirror symmetry of the layering),
            # against the largest field
            assert per_ch[0] < 1e-6 and per_ch[2] < 1e-6 and overall < 1e-6, (isrc, f, per_ch, overall)
    finally:
        ses.close()
