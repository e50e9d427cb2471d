"""Write tests/golden/configs_oracle.npz: full-size solves by the oracle.

For each case the oracle (oracle/csem_oracle.c, itself pinned to the
reference build and to the mathematics) runs the reference's per-frequency
unit (main.c:93-123: inject_source, fbicgstab(fdfd_apply, lusgs_apply) at
tol = (float)tol / f, extract_emf) from x = 0, once in the reference's
serial summation order (cdotprod.c:18) and once in the device's chunked
order.  Stored per (case, frequency, order): the iteration index at
convergence, the relative-residual history, every receiver value in double
precision before the float32 store (Ex, Ey, Ez, plus Hx, Hy, Hz where the
case asks for them) and the SHA-256 of the solution vector's bytes, so a
device solve can be checked bit for bit without storing the vector.

Cases:
  run_layered  the reference's own run_layered example (100^3, dz=50,
               nb=10, airbc=1, 1.25 Hz, 641 receivers), from
               tests/golden/run_layered.npz
  layered_64   64^3 water + 1 ohm-m + 100 ohm-m slab, 0.25 Hz, lextend 40 km
  layered_100  100^3 with 8 seeded layers, 0.25 / 0.75 / 1.25 Hz (budget
               2000 iterations: 0.25 Hz needs more than 1000)
  vti_marine   180 x 120 x 120 anisotropic model (11.1 M unknowns), 1.25 and
               0.75 Hz of its six, six channels; F-BiCGStab in both orders
               at 1.25 Hz and the chunked order at 0.75 Hz, and FGMRES(10)
               in the chunked order at 1.25 Hz (key suffix _fgmres)
  topo_marine  125 x 125 x 100 model with seafloor topography, its first
               source, 1.25 Hz, chunked order

    python tests/golden/make_configs.py [-j 8]

takes about half an hour on 8 cores (the 0.25 Hz 100^3 solve is the
longest).  The data it writes is small (receivers and hashes only).
"""
import argparse
import hashlib
import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from demo_matrixfree_csem_amd import surveys  # noqa: E402
import oracle as O  # noqa: E402

OUT = os.path.join(HERE, "configs_oracle.npz")


def run_layered_survey():
    """The run_layered inputs as a Survey (laterally uniform profiles
    expanded to the 100^3 grid; sources and receivers from its tables)."""
    g = np.load(os.path.join(HERE, "run_layered.npz"))
    nx, ny, nz = len(g["fx1"]) - 1, len(g["fx2"]) - 1, len(g["fx3"]) - 1
    rho = [np.ascontiguousarray(np.broadcast_to(g[k + "_profile"][:, None, None], (nz, ny, nx)), np.float32)
           for k in ("frho11", "frho22", "frho33")]
    tab = lambda t: np.array([[float(v) for v in l.split()[:5]] for l in str(t).splitlines()[1:] if l.strip()],
                             np.float32)
    src, rec = tab(g["sources_txt"]), tab(g["receivers_txt"])
    return surveys.Survey("run_layered", surveys.Model(g["fx1"], g["fx2"], g["fx3"], *rho),
                          surveys.Acquisition(src, rec), [1.25],
                          dict(fvm=0, airbc=1, preco=1, algopt=1, tol=1e-8, niter=1000, nb=10,
                               chsrc="Ex", chrec="Ex,Ey,Ez"))


def cases():
    l100 = surveys.layered_100()
    l100.flags["niter"] = 2000
    vti = surveys.vti_marine()
    vti.freqs = [1.25, 0.75]
    topo = surveys.topo_marine()
    topo.freqs = [1.25]
    return {"run_layered": run_layered_survey(), "layered_64": surveys.layered_64(), "layered_100": l100,
            "vti_marine": vti, "topo_marine": topo}


# (frequency index, summation order, solver) per case; the default is every
# frequency in both orders with F-BiCGStab
TASKS = {
    "vti_marine": [(0, 0, "bicgstab"), (0, 1, "bicgstab"), (1, 1, "bicgstab"), (0, 1, "fgmres")],
    "topo_marine": [(0, 1, "bicgstab")],
}


CHANNELS = {"Ex": 0, "Ey": 1, "Ez": 2, "Hx": 3, "Hy": 4, "Hz": 5}


def oracle_for(s, nfft=1024):
    m = s.model
    return O.Oracle(m.x1node, m.x2node, m.x3node, m.rho11, m.rho22, m.rho33, nb=s.flags["nb"],
                    lextend=float(s.flags.get("lextend", -1.0)), airbc=s.flags["airbc"],
                    fvm=s.flags.get("fvm", 0), nfft=nfft)


def solve(task):
    name, ifreq, red, algo = task
    s = cases()[name]
    orc = oracle_for(s)
    f = s.freqs[ifreq]
    orc.set_freq(f)
    orc.set_red(red)
    tol = float(np.float32(np.float32(s.flags["tol"]) / np.float32(f)))
    src = s.acq.sources[0]
    b = orc.inject(0, float(src[0]), float(src[1]), float(src[2]))
    if algo == "fgmres":
        m = int(s.flags.get("nrestart", 10))
        x, it, hist = orc.fgmres(b, int(s.flags["niter"]) // m, tol, m)
    else:
        x, it, hist = orc.fbicgstab(b, int(s.flags["niter"]), tol)
    chans = [CHANNELS[c] for c in s.flags["chrec"].split(",")]
    rec = np.array([[orc.extract(ch, float(r[0]), float(r[1]), float(r[2]), x) for r in s.acq.receivers]
                    for ch in chans])
    print("%s f=%g red=%d %s iter=%d" % (name, f, red, algo, it), flush=True)
    return task, dict(iter=it, hist=hist, rec=rec, sha=hashlib.sha256(x.tobytes()).hexdigest(), tol=tol)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("--only", default=None, help="comma list of case names")
    a = ap.parse_args()
    cs = cases()
    names = a.only.split(",") if a.only else list(cs)
    tasks = []
    for n in names:
        tl = TASKS.get(n) or [(i, red, "bicgstab") for i in range(len(cs[n].freqs))
                              for red in (O.RED_SERIAL, O.RED_CHUNKED)]
        tasks += [(n, i, red, algo) for i, red, algo in tl]
    # longest first
    tasks.sort(key=lambda t: -np.prod(cs[t[0]].model.shape) / cs[t[0]].freqs[t[1]])
    out = dict(np.load(OUT)) if os.path.exists(OUT) else {}
    with Pool(a.j) as pool:
        for (n, i, red, algo), r in pool.imap_unordered(solve, tasks):
            key = "%s_f%d_%s%s" % (n, i, "serial" if red == O.RED_SERIAL else "chunked",
                                   "_fgmres" if algo == "fgmres" else "")
            out[key + "_iter"] = np.int64(r["iter"])
            out[key + "_hist"] = r["hist"]
            out[key + "_rec"] = r["rec"]
            out[key + "_sha256"] = np.array(r["sha"])
            out[key + "_tol"] = np.float64(r["tol"])
            np.savez_compressed(OUT, **out)


if __name__ == "__main__":
    main()
