import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("fdtd-cuda_amd", "oracle", ""):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def small_params(mod, n=24, ncpml=6, radius=8e-9, steps=160, nfreq=8, block=8, sphere=1, phi=0.0, exact=0):
    """Small grid for parity runs (T = n + 2*ncpml + 1), short pulse so fields are large."""
    p = mod.default_params()
    p.NCELLX = p.NCELLY = p.NCELLZ = n
    p.NcpmlX = p.NcpmlY = p.NcpmlZ = ncpml
    p.tfsf_inset = 4
    p.sca_offset = 2
    p.abs_offset = 2
    p.nano_sphere = sphere
    p.nano_sphere_radius = radius
    p.BandWidth = 2.5e16          # width ~20 steps, delay ~120 steps at dx = 2 nm
    p.Tend = steps
    p.NUM_freq = nfreq
    p.lam_min, p.lam_max = 400e-9, 900e-9
    p.dft_block = block
    p.inc_phi = phi
    p.dft_exact = exact
    return p
