"""Write the golden fixtures under tests/golden/.

Two sources, both data only (inputs and expected outputs):

1. run_layered.npz -- the reference's checked-in run_layered example:
   node files fx1..fx3, the (laterally uniform) resistivity profiles of
   frho11..frho33, the source/receiver tables, the solver output
   emf_0001.txt (1.25 Hz, 641 receivers x Ex,Ey,Ez, float32 printed %e),
   the convergence history iterate_fbicgstab.txt (166 iterations) and the
   extended-grid dumps fx*_extended / frho*_extended (profiles).
2. ref_tiny.npz -- outputs of the reference's own routines, compiled from
   its sources by `make -C oracle ref`, on small seeded models from
   demo_matrixfree_csem_amd.surveys.tiny (grid arrays, edge conductivities,
   fdfd_apply, lusgs_apply, inject_source, extract_emf and full solves).
3. ref_acsr.npz -- the same build's matrix-storage path (Acsr.c): the
   assembled CSR arrays and Acsr_apply / Acsr_lusgs_apply of a vector.

Run in a container that holds /root/reference:
    python tests/golden/make_golden.py [run_layered] [ref_tiny] [ref_acsr]
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from demo_matrixfree_csem_amd import surveys  # noqa: E402
import oracle as O  # noqa: E402

REF = os.environ.get("CSEM_REFERENCE", "/root/reference")


def _profile(path, shape):
    a = np.fromfile(path, np.float64 if path.endswith("_extended") else np.float32).reshape(shape)
    assert np.all(a == a[:, :1, :1]), path + " is not laterally uniform"
    return a[:, 0, 0].copy()


def run_layered():
    d = os.path.join(REF, "run_layered")
    out = {}
    for k in ("fx1", "fx2", "fx3"):
        out[k] = np.fromfile(os.path.join(d, k), np.float32)
    nz, ny, nx = len(out["fx3"]) - 1, len(out["fx2"]) - 1, len(out["fx1"]) - 1
    for k in ("frho11", "frho22", "frho33"):
        out[k + "_profile"] = _profile(os.path.join(d, k), (nz, ny, nx))
    for k in ("fx1_extended", "fx2_extended", "fx3_extended"):
        out[k] = np.fromfile(os.path.join(d, k), np.float64)
    n1, n2, n3 = len(out["fx1_extended"]) - 1, len(out["fx2_extended"]) - 1, len(out["fx3_extended"]) - 1
    for k in ("frho11_extended", "frho22_extended", "frho33_extended"):
        out[k + "_profile"] = _profile(os.path.join(d, k), (n3, n2, n1))
    for k in ("sources.txt", "receivers.txt", "emf_0001.txt", "iterate_fbicgstab.txt"):
        with open(os.path.join(d, k)) as f:
            out[k.replace(".", "_")] = np.array(f.read())
    np.savez_compressed(os.path.join(HERE, "run_layered.npz"), **out)


def _rand(n, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal(n) + 1j * r.standard_normal(n)


CASES = [  # name, survey kwargs, flag overrides
    ("a0", dict(seed=5, nx=6, ny=6, nz=5, airbc=0), dict(nb=2, rho_air=3.0)),
    ("a1", dict(seed=6, nx=6, ny=6, nz=5, airbc=1), dict(nb=2)),
    ("f1", dict(seed=7, nx=6, ny=6, nz=5, airbc=1), dict(nb=2, fvm=1)),
]


def ref_tiny():
    out = {}
    for name, kw, fl in CASES:
        s = surveys.tiny(**kw)
        s.flags.update(fl)
        d = tempfile.mkdtemp()
        surveys.write_files(d, s)
        argv = s.argv(d)
        O.run_ref(d, argv, "grid")
        for a in ("x1", "x2", "x3", "x1s", "x2s", "x3s", "d1", "d2", "d3", "d1s", "d2s", "d3s",
                  "sigma11", "sigma22", "sigma33"):
            out["%s_%s" % (name, a)] = np.fromfile(os.path.join(d, "ref_" + a), np.float64)
        info = np.fromfile(os.path.join(d, "ref_info"), np.float64)
        out[name + "_info"] = info
        lE = int(info[6])
        f = 0.75
        x = _rand(lE, 100)
        x.tofile(os.path.join(d, "xin"))
        if not s.flags["airbc"]:
            O.run_ref(d, argv, "apply", hfreq=f, hin=os.path.join(d, "xin"))
            out[name + "_apply"] = np.fromfile(os.path.join(d, "ref_y"), np.complex128)
        O.run_ref(d, argv, "lusgs", hfreq=f, hin=os.path.join(d, "xin"))
        out[name + "_lusgs"] = np.fromfile(os.path.join(d, "ref_y"), np.complex128)
        # the other preconditioners: A-phi (preco=3) on every case, inner
        # GMRES (preco=2, mp=8) where the operator runs (airbc=0)
        O.run_ref(d, argv + ["preco=3"], "aphi", hfreq=f, hin=os.path.join(d, "xin"))
        out[name + "_aphi"] = np.fromfile(os.path.join(d, "ref_y"), np.complex128)
        if not s.flags["airbc"]:
            O.run_ref(d, argv + ["mp=8"], "gmresprec", hfreq=f, hin=os.path.join(d, "xin"))
            out[name + "_gmresprec"] = np.fromfile(os.path.join(d, "ref_y"), np.complex128)
        O.run_ref(d, argv, "inject", hfreq=f)
        out[name + "_inject"] = np.fromfile(os.path.join(d, "ref_b"), np.complex128)
        O.run_ref(d, argv, "extract", hfreq=f, hin=os.path.join(d, "xin"))
        out[name + "_extract"] = np.stack([np.fromfile(os.path.join(d, "ref_rec%d" % c), np.complex64)
                                           for c in range(3)])
        if not s.flags["airbc"]:
            for algopt, preco in ((1, 1), (1, 0), (2, 1), (2, 0), (1, 2), (2, 2), (1, 3), (2, 3)):
                argv2 = argv + ["preco=%d" % preco, "algopt=%d" % algopt, "niter=60", "nrestart=10",
                                "freqs=0.75", "verb=1", "mp=8"]
                stdout = O.run_ref(d, argv2, "solve")
                tag = "%s_solve_p%d_a%d" % (name, preco, algopt)
                out[tag] = np.fromfile(os.path.join(d, "ref_x0"), np.complex128)
                with open(os.path.join(d, "emf_0001.txt")) as fh:
                    out[tag + "_emf"] = np.array(fh.read())
                out[tag + "_stdout"] = np.array(stdout[-4000:])
    np.savez_compressed(os.path.join(HERE, "ref_tiny.npz"), **out)


ACSR_CASES = [  # name, survey kwargs, flag overrides, frequency (airbc=0: Acsr.c assembles the explicit-air grid)
    ("c0", dict(seed=8, nx=5, ny=4, nz=3, airbc=0), dict(nb=2, rho_air=3.0), 0.75),
    ("c1", dict(seed=9, nx=4, ny=5, nz=4, airbc=0), dict(nb=2, rho_air=10.0), 1.25),
]


def ref_acsr():
    """ref_acsr.npz: the reference's Acsr_init arrays (row_ptr, col_ind,
    val as Acsr.c builds them) and its Acsr_apply / Acsr_lusgs_apply of a
    seeded vector, on small explicit-air models."""
    out = {}
    for name, kw, fl, f in ACSR_CASES:
        s = surveys.tiny(**kw)
        s.flags.update(fl)
        s.freqs = [f]
        d = tempfile.mkdtemp()
        surveys.write_files(d, s)
        argv = s.argv(d)
        O.run_ref(d, argv, "grid")
        lE = int(np.fromfile(os.path.join(d, "ref_info"), np.float64)[6])
        x = _rand(lE, 200)
        x.tofile(os.path.join(d, "xin"))
        O.run_ref(d, argv, "acsr", hin=os.path.join(d, "xin"))
        out[name + "_x"] = x
        out[name + "_row_ptr"] = np.fromfile(os.path.join(d, "ref_row_ptr"), np.int32)
        out[name + "_col"] = np.fromfile(os.path.join(d, "ref_col"), np.int32)
        out[name + "_val"] = np.fromfile(os.path.join(d, "ref_val"), np.complex128)
        out[name + "_apply"] = np.fromfile(os.path.join(d, "ref_y"), np.complex128)
        out[name + "_sgs"] = np.fromfile(os.path.join(d, "ref_ysgs"), np.complex128)
    np.savez_compressed(os.path.join(HERE, "ref_acsr.npz"), **out)


if __name__ == "__main__":
    if not os.path.isdir(REF):
        sys.exit("reference not found at %s" % REF)
    which = sys.argv[1:] or ["run_layered", "ref_tiny", "ref_acsr"]
    for w in which:
        globals()[w]()
        print(w + ".npz", os.path.getsize(os.path.join(HERE, w + ".npz")))
