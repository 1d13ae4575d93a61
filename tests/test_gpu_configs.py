"""The benchmark configurations at sizes the FP64 oracle finishes in seconds,
through the C ABI.

* The 64^3 plain-method case in full: 64^3 interior at dx = 2 nm, 10-cell CPML
  (m = 3, kappa_max = 5, alpha_max = 0.05), TFSF 8 cells inside the CPML,
  x-polarised plane wave along z (600 nm centre, BandWidth 6.79e14 Hz for a
  300 nm FWHM), 20 nm gold sphere, 1000 steps, 16 wavelengths 400-900 nm,
  scattering box 4 cells outside the TFSF box.  Spectra within 1e-4 of their
  maximum, fields within 2e-4 of the largest E or H value.
* The hydrodynamic case down-scaled to 96^3 interior at dx = 1 nm (10 nm
  sphere = 10 cells, beta = 1.27e6 m/s, 48 wavelengths, a short pulse), 700
  steps: fields and face fluxes against the oracle.
* The adjoint gradient against central finite differences of the GPU's own
  forward objective (h = 1e-2 on four voxels).
"""
import numpy as np
import pytest

import fdtd_amd as G
import oracle as O

pytestmark = pytest.mark.gpu


def plain_method_params(mod):
    p = mod.default_params()
    p.NCELLX = p.NCELLY = p.NCELLZ = 64
    p.NcpmlX = p.NcpmlY = p.NcpmlZ = 10
    p.dx = p.dy = p.dz = 2e-9
    p.cpml_exp, p.kappa_max, p.alpha_max = 3, 5.0, 0.05
    p.tfsf_inset = 8
    p.sca_offset = 4
    p.abs_offset = 5
    p.inc_phi = 90.0
    p.TimeFactor = 3 ** 0.5
    p.lam_0 = 600e-9
    p.BandWidth = 6.79e14
    p.nano_sphere = 1
    p.nano_sphere_radius = 20e-9
    p.material = 3
    p.Tend = 1000
    p.NUM_freq = 16
    p.lam_min, p.lam_max = 400e-9, 900e-9
    p.dft_block = 64
    return p


def field_errors(s, o):
    fo = [o.get_field(c) for c in range(6)]
    fs = [s.get_field(c) for c in range(6)]
    me = max(np.abs(f).max() for f in fo[:3])
    mh = max(np.abs(f).max() for f in fo[3:])
    return [float(np.abs(fs[c] - fo[c]).max() / (me if c < 3 else mh)) for c in range(6)]


def spec_err(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_plain_method_64_full_run():
    p = plain_method_params(O)
    s = G.Simulation(p)
    o = O.Oracle(p)
    assert s.get_info().n_dispersive[0] == o.get_info().n_dispersive[0] > 0
    s.step(1000)
    o.step(1000)
    info = s.get_info()
    print("pulse width %.0f, delay %.0f steps" % (info.width, info.delay), flush=True)
    assert np.abs(o.get_field(0)).max() > 0
    errs = field_errors(s, o)
    fs, fo = s.flux(), o.flux()
    e = {k: spec_err(fs[k], fo[k]) for k in ("sca", "abs", "ext")}
    print("64^3, 1000 steps: fields", ["%.1e" % x for x in errs], "spectra", {k: "%.1e" % v for k, v in e.items()},
          "ext", fo["ext"][:4], flush=True)
    assert max(errs) < 2e-4
    assert max(e.values()) < 1e-4


def test_hydrodynamic_96_downscaled():
    p = O.default_params()
    p.NCELLX = p.NCELLY = p.NCELLZ = 96
    p.NcpmlX = p.NcpmlY = p.NcpmlZ = 12
    p.dx = p.dy = p.dz = 1e-9
    p.TimeFactor = 3 ** 0.5
    p.inc_phi = 90.0
    p.nano_sphere = 1
    p.nano_sphere_radius = 10e-9
    p.material = 3
    p.NONLOCAL = 3
    p.hydro_beta = 1.27e6
    p.tfsf_inset = 10
    p.BandWidth = 2.5e16          # short pulse (width ~40 steps at dx = 1 nm) so it crosses the sphere in the run
    p.lam_0 = 500e-9
    p.NUM_freq = 48
    p.lam_min, p.lam_max = 300e-9, 700e-9
    p.dft_block = 64
    steps = 700
    s = G.Simulation(p)
    o = O.Oracle(p)
    assert s.get_info().hydro == 1 and s.get_info().n_dispersive[0] > 0
    s.step(steps)
    o.step(steps)
    errs = field_errors(s, o)
    raw_s, raw_o = s.flux()["raw"], o.flux()["raw"]
    err_raw = spec_err(raw_s, raw_o)
    print("hydro 96^3, %d steps: fields" % steps, ["%.1e" % x for x in errs], "face fluxes %.1e" % err_raw, flush=True)
    assert max(errs) < 2e-4
    assert err_raw < 1e-4


def test_adjoint_gradient_against_finite_differences():
    import test_gpu_adjoint as A
    p, rho, lo, r0 = A.problem(G)
    freqs = A.C0 / A.LAMS
    sim = G.Simulation(p)
    sim.set_design(lo, rho)
    J, g = sim.adjoint(A.STEPS, r0, freqs, checkpoint=0)

    def objective(r):
        s = G.Simulation(p)
        s.set_design(lo, r)
        s.set_point_monitor(r0, A.STEPS)
        s.step(A.STEPS)
        e2, _ = s.point_spectrum(freqs)
        return float(np.sum(e2))

    assert objective(rho) == pytest.approx(J, rel=1e-12)
    h = 1e-2
    rng = np.random.default_rng(5)
    flat = [tuple(int(v) for v in np.unravel_index(q, rho.shape)) for q in rng.choice(rho.size, 4, replace=False)]
    gmax = np.abs(g).max()
    for v in flat:
        rp, rm = rho.copy(), rho.copy()
        rp[v] += h
        rm[v] -= h
        fd = (objective(rp) - objective(rm)) / (2 * h)
        print("voxel", v, "adjoint %.6e fd %.6e rel-to-max %.2e" % (g[v], fd, abs(fd - g[v]) / gmax), flush=True)
        assert abs(fd - g[v]) < 1e-3 * gmax
