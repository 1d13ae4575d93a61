"""Periodic x-y mode (Periodic_XY) and its reflectance spectra on the GPU
against the FP64 oracle, through the C ABI.

* A metal film across a periodic column, fields after 600 steps (split and
  fused step), for the Drude + 2 CP gold and the Lorentz-Drude material;
  reflectance and transmittance after 8000 steps against the oracle's (1e-4
  absolute; both are fractions of the incident power) and against the Airy
  formula (the oracle's own pin).
* A staircased sphere in a periodic 26 x 26 cell (a square array of
  particles), fields against the oracle.
* Vacuum: R = 0 and T = 1; non-periodic handles reject fdtd_reflectance.
"""
import numpy as np
import pytest

import fdtd_amd
import oracle as O
from fdtd_amd.abi import FDTD_ERR_INVALID

pytestmark = pytest.mark.gpu

TOL = 2e-4


def params(mod, material=3, n_col=4, nz=160, steps=8000, dz=1e-9):
    p = mod.default_params()
    p.periodic_xy = 1
    p.NcpmlX = p.NcpmlY = 0
    p.NcpmlZ = 12
    p.NCELLX = p.NCELLY = n_col
    p.NCELLZ = nz
    p.dx = p.dy = p.dz = dz
    p.nano_sphere = 0
    p.material = material
    p.TimeFactor = 0.9
    p.inc_phi = 90.0
    p.BandWidth = 1.5e15
    p.NUM_freq = 13
    p.lam_min, p.lam_max = 400e-9, 1000e-9
    p.Tend = steps
    p.dft_block = 16
    return p


def field_errors(s, o):
    fo = [o.get_field(c) for c in range(6)]
    fs = [s.get_field(c) for c in range(6)]
    me = max(np.abs(f).max() for f in fo[:3])
    mh = max(np.abs(f).max() for f in fo[3:])
    return [float(np.abs(fs[c] - fo[c]).max() / (me if c < 3 else mh)) for c in range(6)]


def film_pair(material, steps, fused=True):
    s = fdtd_amd.Simulation(params(fdtd_amd, material, steps=steps))
    o = O.Oracle(params(O, material, steps=steps))
    k0 = s.info.tfsf[2] + 30
    s.add_layer(k0, k0 + 20)
    o.add_layer(k0, k0 + 20)
    assert list(s.get_info().n_dispersive) == list(o.get_info().n_dispersive) == [16 * 20] * 3
    o.step(steps)
    if fused:
        s.step(steps)
    else:
        s.split_step(steps)
    return s, o


@pytest.mark.parametrize("material", [3, 7])
@pytest.mark.parametrize("fused", [True, False])
def test_periodic_film_fields(material, fused):
    s, o = film_pair(material, 600, fused)
    errs = field_errors(s, o)
    print(f"film material {material} fused={fused}: {['%.1e' % e for e in errs]}", flush=True)
    assert max(errs) < TOL


@pytest.mark.parametrize("material", [3, 7])
def test_periodic_film_reflectance(material):
    s, o = film_pair(material, 8000)
    rs, ro = s.reflectance(), o.reflectance()
    d = {k: float(np.abs(rs[k] - ro[k]).max()) for k in ("R", "T")}
    d.update({k: float(np.abs(rs[k] - ro[k]).max() / np.abs(ro[k]).max()) for k in ("ref_r", "ref_t")})
    print(f"film material {material}: R {np.round(rs['R'][::4], 4)} T {np.round(rs['T'][::4], 4)}; "
          f"GPU - oracle {d}", flush=True)
    assert max(d.values()) < 1e-4


def test_periodic_vacuum_and_errors():
    s = fdtd_amd.Simulation(params(fdtd_amd, steps=4000))
    s.step(4000)
    r = s.reflectance()
    assert np.abs(r["R"]).max() < 1e-5 and np.abs(r["T"] - 1).max() < 2e-3
    q = fdtd_amd.Simulation(params(fdtd_amd, steps=10))
    with pytest.raises(fdtd_amd.FDTDError):
        q.add_layer(2, 5)              # inside the CPML / before the source plane
    p = fdtd_amd.default_params()
    t = fdtd_amd.Simulation(p)
    with pytest.raises(fdtd_amd.FDTDError) as e:
        t.reflectance()
    assert e.value.code == FDTD_ERR_INVALID


@pytest.mark.parametrize("fused", [True, False])
def test_periodic_sphere_array(fused):
    """A 12 nm gold sphere in a 26 x 26 periodic cell (dx = 2 nm): the sphere's
    near field wraps across the x and y edges."""
    def mk(mod):
        p = params(mod, 3, n_col=26, nz=60, steps=500, dz=2e-9)
        p.TimeFactor = 3 ** 0.5
        p.BandWidth = 2.5e16
        return p
    s = fdtd_amd.Simulation(mk(fdtd_amd))
    o = O.Oracle(mk(O))
    c = (13, 13, s.info.tfsf[2] + 16)
    s.add_sphere(c, 12.5e-9)
    o.add_sphere(c, 12.5e-9)
    assert list(s.get_info().n_dispersive) == list(o.get_info().n_dispersive)
    o.step(500)
    if fused:
        s.step(500)
    else:
        s.split_step(500)
    errs = field_errors(s, o)
    rs, ro = s.reflectance(), o.reflectance()
    print(f"sphere array fused={fused}: {['%.1e' % e for e in errs]}, R diff {np.abs(rs['R'] - ro['R']).max():.1e}",
          flush=True)
    assert max(errs) < TOL
