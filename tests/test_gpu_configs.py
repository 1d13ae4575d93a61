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
from fdtd_amd import workloads as W

pytestmark = pytest.mark.gpu


def plain_method_params(mod):
    return W.plain_method(mod)


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


# ---------------------------------------------------------------- the 32-rod sweep

def test_sweep_32_seeded_rods_batched_equals_alone():
    """The sweep workload itself: the 32 seeded nanorod draws (edges 96-192,
    dx = 1.5 nm) advanced in one batch with the simulations' own tiles equal
    each simulation stepped alone, bit for bit, over 64 steps (graph replay of
    16 sampling periods)."""
    from fdtd_amd import sweep
    cfgs, redraws = sweep.draw(20240601)
    assert len(cfgs) == 32 and redraws == 3
    a = [sweep.make_sim(G.Simulation, G, c) for c in cfgs]
    b = G.Batch(a, flags=G.BATCH_NO_TUNE)
    b.run(64)
    b.close()
    for i, c in enumerate(cfgs):
        s = sweep.make_sim(G.Simulation, G, c)
        s.step(64)
        for comp in range(6):
            assert np.array_equal(a[i].get_field(comp), s.get_field(comp)), (i, comp)
        assert np.array_equal(a[i].flux()["raw"], s.flux()["raw"]), i
        s.close()
        a[i].close()
        print("draw", i, c["edges"], "bit-exact", flush=True)


@pytest.mark.parametrize("which", ["smallest", "largest"])
def test_sweep_extreme_draws_against_the_oracle(which):
    """The smallest and the largest grid of the 32 draws (their rods, CPML,
    monitors and polarisation), batched with the autotuned tiles, against the
    FP64 oracle; a short pulse (BandWidth 2.5e16 Hz) so the fields are large
    within the 300 steps the oracle runs."""
    from fdtd_amd import sweep
    cfgs, _ = sweep.draw(20240601)
    cells = [np.prod(c["edges"]) for c in cfgs]
    cfg = cfgs[int(np.argmin(cells) if which == "smallest" else np.argmax(cells))]
    steps = 300

    def params(mod):
        p = sweep.params_for(mod, cfg, steps)
        p.BandWidth = 2.5e16
        p.lam_0 = 600e-9
        return p

    s = G.Simulation(params(G))
    c = [n // 2 for n in s.shape]
    s.add_rod(c, cfg["axis"], cfg["length"], cfg["diameter"] / 2)
    b = G.Batch([s])
    b.run(steps)
    b.close()
    o = O.Oracle(params(O))
    o.set_materials(s.get_materials(1), s.get_materials(2), s.get_materials(3))
    o.step(steps)
    errs = field_errors(s, o)
    fs, fo = s.flux(), o.flux()
    e = spec_err(fs["raw"], fo["raw"])
    print(which, cfg["edges"], "fields", ["%.1e" % x for x in errs], "face fluxes %.1e" % e, flush=True)
    assert s.get_info().n_dispersive[0] > 0
    assert max(errs) < 2e-4
    assert e < 1e-4


# ---------------------------------------------------------------- inverse design at full size

@pytest.mark.timeout(900)
def test_inverse_design_gradient_at_full_size():
    """The inverse-design workload as configured: 256^3 interior at dx = 2 nm
    (281^3 cells), a 64x64x32 density block (rho ~ U[0, 1), seeded), |E|^2 at
    633/785/850 nm 5 nm above the block, 10,000 steps with a checkpoint every
    500 and segment recompute.  dJ/drho against central finite differences of
    the GPU's own objective with h = 1e-3 on 16 seeded voxels, within 1e-3 of
    the largest gradient entry."""
    p = W.design(G)
    T = 256 + 25
    lo, r0 = W.design_geometry(T)
    rho = np.random.default_rng(7).uniform(0.0, 1.0, size=W.DESIGN_BLOCK).astype(np.float32)
    freqs = np.array(W.design_freqs())
    steps = W.DESIGN_STEPS
    sim = G.Simulation(p)
    sim.set_design(lo, rho)
    J, g = sim.adjoint(steps, r0, freqs, checkpoint=W.DESIGN_CHECKPOINT)
    sim.close()
    print("J %.6e |g| %.6e max|g| %.6e" % (J, np.linalg.norm(g), np.abs(g).max()), flush=True)
    assert J > 0 and np.all(np.isfinite(g))

    def objective(r):
        s = G.Simulation(p)
        s.set_design(lo, r)
        s.set_point_monitor(r0, steps)
        s.step(steps)
        e2, _ = s.point_spectrum(freqs)
        s.close()
        return float(np.sum(e2))

    assert objective(rho) == pytest.approx(J, rel=1e-12)
    h = np.float32(1e-3)
    rng = np.random.default_rng(11)
    flat = [tuple(int(v) for v in np.unravel_index(q, rho.shape)) for q in rng.choice(rho.size, 16, replace=False)]
    gmax = np.abs(g).max()
    worst = 0.0
    for v in flat:
        rp, rm = rho.copy(), rho.copy()
        rp[v] = min(rho[v] + h, 1.0)
        rm[v] = max(rho[v] - h, 0.0)
        fd = (objective(rp) - objective(rm)) / (float(rp[v]) - float(rm[v]))
        err = abs(fd - g[v]) / gmax
        worst = max(worst, err)
        print("voxel", v, "adjoint %.6e fd %.6e rel-to-max %.2e" % (g[v], fd, err), flush=True)
    assert worst < 1e-3
