"""Writes the FP64 golden fixtures of the long benchmark runs.

Run from the repository root after `make -C oracle`:

    python tests/golden/make_fixtures.py dimer     # 281^3 dimer, 20,000 steps (~45 min on 8 cores)
    python tests/golden/make_fixtures.py hydro96   # 96^3 hydrodynamic copy, 40,000 steps

Each fixture is an .npz of plain arrays (no pickled objects) holding the FP64
oracle's spectra after the full run: per-frequency scattering, absorption and
extinction cross sections, the 12 raw face fluxes, the source spectrum, and
(dimer) the gap-field sums sum_t E_c^{t+1}(r) e^{i 2 pi f (t+1) dt} at the gap
centre, plus the six field components on two lines through the grid centre.
The oracle is the FP64 restatement of the reference's update equations; these
files let the GPU tests compare a full-length FP32 run with it in seconds.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "fdtd-cuda_amd"), os.path.join(ROOT, "oracle")]

import oracle as O  # noqa: E402
from fdtd_amd import workloads as W  # noqa: E402


def run(o, steps, chunk=500):
    t0 = time.time()
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        o.step(n)
        done += n
        el = time.time() - t0
        print(f"  step {done}/{steps}  {el:.0f} s  eta {el / done * (steps - done):.0f} s", flush=True)
    return time.time() - t0


def lines(o, c):
    """Six components on the x line (j = k = c) and the z line (i = j = c)."""
    out = {}
    for comp in range(6):
        f = o.get_field(comp)
        out[f"xline_{comp}"] = f[:, c[1], c[2]].copy()
        out[f"zline_{comp}"] = f[c[0], c[1], :].copy()
        del f
    return out


def spectra(o):
    fl = o.flux()
    freq, einc = o.spectrum_setup()
    return {"freq": freq, "e_incident": einc, "sca": fl["sca"], "abs": fl["abs"], "ext": fl["ext"], "raw": fl["raw"]}


def dimer(path):
    p = W.dimer(O)
    o = O.Oracle(p)
    T = o.info.NX
    for c in W.dimer_centres(T):
        o.add_sphere(c, W.DIMER_RADIUS)
    gap = W.dimer_gap(T)
    freq, _ = o.spectrum_setup()
    o.set_objective(gap, freq)
    print(f"dimer {T}^3, {p.Tend} steps, dispersive cells {list(o.info.n_dispersive)}, {o.threads()} threads",
          flush=True)
    wall = run(o, p.Tend)
    res = spectra(o)
    res["gap_terms"] = o.objective_terms()
    res["gap"] = np.array(gap)
    res["n_dispersive"] = np.array(list(o.get_info().n_dispersive))
    res["steps"] = np.array(p.Tend)
    res["wall_s"] = np.array(wall)
    res.update(lines(o, gap))
    np.savez_compressed(path, **res)
    print("wrote", path, flush=True)


def hydro96(path):
    p = W.hydro(O, n=96, dx=1e-9)
    o = O.Oracle(p)
    T = o.info.NX
    print(f"hydro {T}^3, {p.Tend} steps, dispersive cells {list(o.info.n_dispersive)}, {o.threads()} threads",
          flush=True)
    wall = run(o, p.Tend, chunk=2000)
    res = spectra(o)
    res["n_dispersive"] = np.array(list(o.get_info().n_dispersive))
    res["steps"] = np.array(p.Tend)
    res["wall_s"] = np.array(wall)
    res.update(lines(o, (T // 2, T // 2, T // 2)))
    np.savez_compressed(path, **res)
    print("wrote", path, flush=True)


if __name__ == "__main__":
    what = sys.argv[1]
    fn = {"dimer": dimer, "hydro96": hydro96}[what]
    fn(os.path.join(HERE, {"dimer": "dimer281_fp64.npz", "hydro96": "hydro96_fp64.npz"}[what]))
