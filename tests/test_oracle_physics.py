"""Running checks of the FP64 oracle against physics it must reproduce
(CPU only).  Each one runs the oracle's update code, so a sign, slab or index
mistake in the restated equations fails it:

* CPML (psi recursions of main.cu:834-1037 and their E counterparts): a
  divergence-free pulse radiated from the grid centre, compared with the same
  pulse on a grid larger by 60 cells per axis over a window that ends before
  the larger grid's boundary echo returns.  The difference is the CPML's
  reflection; with the CPML switched off the same measure is large.
* Drude + 2 critical-point ADE and the flux integral (main.cu:1178-1219,
  Spectral.cu:1162-1168): a staircased gold sphere against the Mie series
  (Bohren and Huffman, ch. 4) with the same permittivity, at two resolutions;
  the cross sections are normalised physically by the DFT of the incident
  field itself.
* Hydrodynamic model (main.cu:1916-2276, 1220-1271): at beta = 0 it is a
  second-order accurate rediscretisation of the local Drude pole, so it must
  approach the local run as dt^2; the charge-density copies must keep their
  total at zero (continuity in flux form, main.cu:1224-1239) and follow the
  charge Gauss's law gives for the same fields.
* Yee update in vacuum (main.cu:762-1044, 1048-1172): the phase a TFSF plane
  wave gathers along z follows the Yee dispersion relation, not the
  continuum one.
"""
import numpy as np
import pytest
from scipy.special import spherical_jn, spherical_yn

import oracle as O
from conftest import small_params
from fdtd_amd import workloads as W
from test_oracle_closed_forms import drude_cp_eps

C0 = 2.99792458e8
EP0 = 8.85418782E-12
E0 = 1.60217662E-19


# ---------------------------------------------------------------- CPML reflection

def _radiate(n_int, steps_list, which, cpml=True):
    """Oracle run with no source (MAX_AMP = 0) from a divergence-free blob at
    the centre: which = "H" sets (hx, hy) = curl of a Gaussian stream function
    along z, "E" sets (ex, ey) the same way.  Returns the six components in the
    21^3 cells around the centre at each requested step."""
    p = O.default_params()
    p.NCELLX = p.NCELLY = p.NCELLZ = n_int
    p.NcpmlX = p.NcpmlY = p.NcpmlZ = 8
    p.nano_sphere = 0
    p.MAX_AMP = 0.0
    p.tfsf_inset, p.abs_offset, p.sca_offset = 4, 2, 2
    if not cpml:
        p.kappa_max, p.sigma_factor, p.alpha_max = 1.0, 0.0, 0.0
    o = O.Oracle(p)
    T = o.shape[0]
    c = T // 2
    x = np.arange(T, dtype=float) - c
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    psi = np.exp(-(X ** 2 + Y ** 2 + (Z / 1.5) ** 2) / (2 * 2.5 ** 2))
    # discrete curl of (0, 0, psi): its Yee divergence vanishes identically
    a = np.zeros_like(psi)
    b = np.zeros_like(psi)
    a[:, :-1, :] = psi[:, 1:, :] - psi[:, :-1, :]
    b[:-1, :, :] = -(psi[1:, :, :] - psi[:-1, :, :])
    if which == "H":
        o.set_field(3, a)
        o.set_field(4, b)
    else:
        o.set_field(0, b)
        o.set_field(1, -a)
    out, done = {}, 0
    win = (slice(c - 10, c + 11),) * 3
    for s in steps_list:
        o.step(s - done)
        done = s
        out[s] = np.stack([o.get_field(k)[win] for k in range(6)])
    return out


@pytest.mark.parametrize("which", ["H", "E"])
def test_cpml_reflection_below_minus_40_db(which):
    steps = [20, 40, 60, 120, 180, 240]
    small = _radiate(30, steps, which)
    large = _radiate(90, steps, which)
    bare = _radiate(30, steps, which, cpml=False)

    def worst(run):
        # E and H separately against their own peak in the window (the pulse
        # passes through it during the first steps)
        out = 0.0
        for f in (slice(0, 3), slice(3, 6)):
            peak = max(np.abs(large[s][f]).max() for s in steps)
            out = max(out, max(np.abs(run[s][f] - large[s][f]).max() for s in steps) / peak)
        return out

    refl, hard = worst(small), worst(bare)
    print(f"{which} blob: CPML reflection {refl:.2e} of the peak, without CPML {hard:.2e}")
    assert refl < 1e-2
    assert hard > 0.1


# ---------------------------------------------------------------- Mie series

def mie_q(eps, x):
    """Extinction and scattering efficiencies of a sphere of relative
    permittivity eps and size parameter x (Bohren and Huffman, eqs. 4.53,
    4.56-4.57, 4.61-4.62)."""
    m = np.sqrt(eps)
    qe = qs = 0.0
    for n in range(1, int(x + 4 * x ** (1 / 3) + 2) + 3):
        jx, jxp = spherical_jn(n, x), spherical_jn(n, x, derivative=True)
        yx, yxp = spherical_yn(n, x), spherical_yn(n, x, derivative=True)
        mx = m * x
        jm, jmp = spherical_jn(n, mx), spherical_jn(n, mx, derivative=True)
        psi, dpsi = x * jx, jx + x * jxp
        xi, dxi = x * (jx + 1j * yx), (jx + 1j * yx) + x * (jxp + 1j * yxp)
        pm, dpm = mx * jm, jm + mx * jmp
        a = (m * pm * dpsi - psi * dpm) / (m * pm * dxi - xi * dpm)
        b = (pm * dpsi - m * psi * dpm) / (pm * dxi - m * xi * dpm)
        qe += (2 * n + 1) * (a + b).real
        qs += (2 * n + 1) * (abs(a) ** 2 + abs(b) ** 2)
    return 2 * qe / x ** 2, 2 * qs / x ** 2


def test_mie_series_closed_form_limits():
    """The Mie routine itself: a perfectly matched sphere does not scatter,
    and a small sphere follows the Rayleigh (quasi-static) limit."""
    qe, qs = mie_q(1.0 + 0j, 0.5)
    assert abs(qe) < 1e-12 and abs(qs) < 1e-12
    eps, x = -2.5 + 1.0j, 0.01
    alpha = (eps - 1) / (eps + 2)
    qe, qs = mie_q(eps, x)
    assert qs == pytest.approx(8 / 3 * x ** 4 * abs(alpha) ** 2, rel=1e-3)
    assert qe == pytest.approx(4 * x * alpha.imag + 8 / 3 * x ** 4 * abs(alpha) ** 2, rel=1e-3)


def _sphere_cross_sections(n_int, dx, steps):
    """Oracle run of a 20 nm gold sphere (the 64^3 plain-method geometry,
    shrunk to n_int^3 interior) with a short pulse, time step near the Courant
    limit; cross sections normalised by the DFT of the incident field at the
    TFSF plane (1D line at m0, the monitors' stride and (t+1) dt phase), so
    the reference's own normalisation constants drop out."""
    p = W.plain_method(O)
    p.NCELLX = p.NCELLY = p.NCELLZ = n_int
    p.dx = p.dy = p.dz = dx
    p.tfsf_inset, p.sca_offset, p.abs_offset = 4, 2, 3
    p.BandWidth = 1.5e15
    p.TimeFactor = 0.9
    p.Tend = steps
    p.NUM_freq = 11
    p.lam_min, p.lam_max = 450e-9, 750e-9
    o = O.Oracle(p)
    info = o.info
    freq, einc_ref = o.spectrum_setup()
    S = np.zeros(len(freq), complex)
    for t in range(0, steps, p.dft_stride):
        o.step(1)
        e, _ = o.get_line()
        S += e[info.m0] * np.exp(2j * np.pi * freq * (t + 1) * info.dt)
        o.step(p.dft_stride - 1)
    fl = o.flux()
    ncell = int((o.get_materials(0) > 1).sum())
    r = (3 * ncell / (4 * np.pi)) ** (1 / 3) * dx            # equal-volume radius of the staircase
    scale = np.abs(einc_ref) ** 2 / np.abs(S) ** 2 * p.nano_sphere_radius ** 2 / r ** 2
    q = {k: fl[k] * scale for k in ("sca", "abs", "ext")}
    lam = C0 / freq
    mie = np.array([mie_q(drude_cp_eps(2 * np.pi * f, conj_convention=False), 2 * np.pi * r / l)
                    for f, l in zip(freq, lam)])
    return lam, q, mie[:, 0], mie[:, 1]


def test_gold_sphere_against_mie_series():
    """Cross sections of the staircased sphere against Mie at two
    resolutions.  A staircased metal surface converges at first order (the
    corners concentrate the field), so the test asks for the signs, the
    resonance position, agreement within 35 % of the Mie peak at 10 cells per
    radius and an error that shrinks when the cells halve."""
    lam, q10, qe, qs = _sphere_cross_sections(44, 2e-9, 10400)
    _, q5, qe5, qs5 = _sphere_cross_sections(24, 4e-9, 5200)
    for i, l in enumerate(lam):
        print("%4.0f nm  Qext %.4f / Mie %.4f (r=5 cells %.4f)  Qsca %.4f / %.4f" % (
            l * 1e9, q10["ext"][i], qe[i], q5["ext"][i], q10["sca"][i], qs[i]))
    assert np.all(q10["sca"] > 0) and np.all(q10["abs"] > 0)
    assert np.allclose(q10["ext"], q10["sca"] + q10["abs"])
    assert np.argmax(q10["ext"]) == np.argmax(qe)
    err10 = np.max(np.abs(q10["ext"] - qe)) / qe.max()
    err5 = np.max(np.abs(q5["ext"] - qe5)) / qe5.max()
    sca10 = np.max(np.abs(q10["sca"] - qs) / qs)
    print(f"max |Qext - Mie| / Mie peak: r=10 cells {err10:.3f}, r=5 cells {err5:.3f}; Qsca {sca10:.3f}")
    assert err10 < 0.35
    assert err10 < 0.6 * err5
    assert sca10 < 0.35


# ---------------------------------------------------------------- hydrodynamic model

def _hydro_pair(time_factor, steps):
    out = []
    for nl, beta in ((0, 0.0), (3, 0.0)):
        p = small_params(O, n=26, ncpml=6, radius=10e-9, steps=steps)
        p.NONLOCAL = nl
        p.hydro_beta = beta
        p.TimeFactor = time_factor
        o = O.Oracle(p)
        o.step(steps)
        out.append([o.get_field(c) for c in range(6)])
    a, b = out
    m = max(np.abs(f).max() for f in a[:3])
    return max(np.abs(a[c] - b[c]).max() for c in range(3)) / m


def test_hydro_at_zero_beta_converges_to_the_local_drude_model():
    """beta = 0 removes the pressure term: the velocity recursion with the
    explicit current is then another second-order discretisation of the same
    Drude pole, so the two runs must agree to O(dt^2) over the same physical
    time."""
    e1 = _hydro_pair(3 ** 0.5, 400)
    e2 = _hydro_pair(2 * 3 ** 0.5, 800)
    print(f"beta=0 hydro vs local Drude: {e1:.2e} at dt, {e2:.2e} at dt/2 (ratio {e1 / e2:.2f})")
    assert e1 < 1e-3
    assert 2.5 < e1 / e2 < 6.0


def test_hydro_charge_total_and_gauss_law():
    p = small_params(O, n=30, ncpml=6, radius=10e-9, steps=700)
    p.NONLOCAL = 3
    p.hydro_beta = 1.27e6
    p.inc_phi = 90.0
    o = O.Oracle(p)
    dx = p.dx
    eps_inf = o.coeffs()["eps_inf"]
    T = o.shape[0]
    c, r = T // 2, 8
    qg, qn = [], []
    for t in (330, 400, 480, 600):
        o.step(t - o.get_info().t)
        ND = [o.get_state(0, k) for k in range(3)]
        for n in ND:
            # flux-form continuity with a hard wall: the total stays zero
            assert abs(n.sum()) < 1e-12 * np.abs(n).sum()
        mats = [o.get_materials(k) for k in (1, 2, 3)]
        D = [np.where(mats[k] > 1, EP0 * eps_inf * o.get_field(k) + o.get_state(1, k), EP0 * o.get_field(k))
             for k in range(3)]
        div = np.zeros_like(D[0])
        div[1:] += D[0][1:] - D[0][:-1]
        div[:, 1:] += D[1][:, 1:] - D[1][:, :-1]
        div[:, :, 1:] += D[2][:, :, 1:] - D[2][:, :, :-1]
        # charge in the half of the particle at x > c: Gauss's law from the
        # fields, and -e ND of the x copy (at the Ex positions)
        half = (slice(c + 1, c + r + 1), slice(c - r, c + r + 1), slice(c - r, c + r + 1))
        qg.append(div[half].sum() * dx * dx)
        nx = ND[0][:, c - r:c + r + 1, c - r:c + r + 1]
        qn.append(-E0 * (nx[c + 1:c + r + 1].sum() + 0.5 * nx[c].sum()) * dx ** 3)
    qg, qn = np.array(qg), np.array(qn)
    ratio = qg / qn
    print("half-particle charge: Gauss", qg, "density", qn, "ratio", ratio)
    # same sign and time course; the reference's centred continuity stencil
    # takes half a cell of the wall flux, so the density holds about half of
    # the Gauss charge
    assert np.all(np.sign(qg) == np.sign(qn))
    assert np.all((ratio > 1.5) & (ratio < 3.5))


# ---------------------------------------------------------------- periodic films (Periodic_XY)

def _film(material, cells, dz=1e-9, steps=8000):
    """Oracle run of a metal film `cells` thick across a periodic 4 x 4
    column (x, y wrap, CPML along z), normal incidence, 13 wavelengths
    400-1000 nm; R and T from the reflection and transmission planes over the
    incident line's own DFT."""
    p = O.default_params()
    p.periodic_xy = 1
    p.NcpmlX = p.NcpmlY = 0
    p.NcpmlZ = 12
    p.NCELLX = p.NCELLY = 4
    p.NCELLZ = 160
    p.dx = p.dy = p.dz = dz
    p.nano_sphere = 0
    p.material = material
    p.TimeFactor = 0.9
    p.inc_phi = 90.0
    p.BandWidth = 1.5e15
    p.NUM_freq = 13
    p.lam_min, p.lam_max = 400e-9, 1000e-9
    p.Tend = steps
    o = O.Oracle(p)
    k0 = o.info.tfsf[2] + 30
    if cells:
        o.add_layer(k0, k0 + cells)
    o.step(steps)
    freq, _ = o.spectrum_setup()
    return freq, o.reflectance()


def airy(eps, freq, d):
    """Reflectance and transmittance of a film of permittivity eps and
    thickness d in vacuum at normal incidence (Born and Wolf, sec. 1.6)."""
    n = np.sqrt(eps)
    k0 = 2 * np.pi * freq / C0
    r12, t12, t23 = (1 - n) / (1 + n), 2 / (1 + n), 2 * n / (n + 1)
    ph = np.exp(2j * k0 * n * d)
    r = (r12 - r12 * ph) / (1 - r12 * r12 * ph)
    t = t12 * t23 * np.exp(1j * k0 * n * d) / (1 - r12 * r12 * ph)
    return np.abs(r) ** 2, np.abs(t) ** 2


def test_periodic_vacuum_transmits_everything():
    freq, r = _film(3, 0)
    assert np.abs(r["R"]).max() < 1e-6
    assert np.abs(r["T"] - 1).max() < 2e-4


@pytest.mark.parametrize("material,tol", [(3, 5e-3), (7, 1e-3)])
def test_metal_film_matches_the_airy_formula(material, tol):
    """A 20 nm film (20 cells) of the Drude + 2 CP gold or the Lorentz-Drude
    material: R and T against the Airy formula with the model's own
    permittivity.  The film is flat, so no staircase error enters; this pins
    the in-grid ADE update, the plane-wave source and the flux sums."""
    from test_oracle_closed_forms import lorentz_drude_eps
    freq, r = _film(material, 20)
    w = 2 * np.pi * freq
    eps = (np.array([drude_cp_eps(x, conj_convention=False) for x in w]) if material == 3 else lorentz_drude_eps(w))
    Ra, Ta = airy(eps, freq, 20e-9)
    dR, dT = np.abs(r["R"] - Ra).max(), np.abs(r["T"] - Ta).max()
    print(f"material {material} film: max |R - Airy| {dR:.2e}, max |T - Airy| {dT:.2e}")
    assert dR < tol and dT < tol
    # the reference's columns: reflected flux positive, its transmitted column carries
    # the opposite sign (both use Re(Ey Hx* - Ex Hy*))
    assert np.all(r["ref_r"] > 0) and np.all(r["ref_t"] < 0)


# ---------------------------------------------------------------- Yee numerical dispersion

def _point_dft(z_off, freqs):
    """Vacuum run at dx = 20 nm (20 to 45 cells per wavelength), x-polarised
    plane wave along z; DFT of E at the column centre, z_off cells after the
    near z face of the TFSF box, over a window that ends after the pulse has
    passed and before the echoes from the ends of the 1D line (first-order
    Mur, 600 cells long at 100 interior cells) come back."""
    p = O.default_params()
    p.NCELLX = p.NCELLY = 12
    p.NCELLZ = 100
    p.NcpmlX = p.NcpmlY = p.NcpmlZ = 8
    p.dx = p.dy = p.dz = 20e-9
    p.nano_sphere = 0
    p.inc_phi = 90.0
    p.tfsf_inset, p.abs_offset, p.sca_offset = 2, 2, 2
    p.lam_0 = 600e-9
    p.BandWidth = 1.5e15
    p.NUM_freq = 4
    o = O.Oracle(p)
    T = o.shape
    r0 = (T[0] // 2, T[1] // 2, o.info.tfsf[2] + z_off)
    o.set_objective(r0, freqs)
    o.step(int(o.info.delay) + 1000)
    return o.objective_terms()[:, 0], o.info.dt


def test_plane_wave_follows_the_yee_dispersion_relation():
    """The phase a plane wave gathers over D cells along z is k D with k from
    the Yee relation sin(w dt/2)/(c dt) = sin(k dz/2)/dz, not the continuum
    w/c: at 20 cells per wavelength and dt = dz/(2 sqrt(3) c) the two differ
    by 0.36 %, 7e-3 rad over 6 cells, and the measured phase must sit within
    1e-5 rad of the Yee value (its amplitude within 1e-5 of unity; measured
    about 1e-9).  A wrong Courant factor, curl sign or component pairing in
    the update fails it."""
    dz = 20e-9
    lam = np.array([400e-9, 500e-9, 700e-9, 900e-9])
    freqs = C0 / lam
    D = 6
    a1, dt = _point_dft(4, freqs)
    a2, _ = _point_dft(4 + D, freqs)
    ratio = a2 / a1
    w = 2 * np.pi * freqs
    k_yee = 2 / dz * np.arcsin(dz / (C0 * dt) * np.sin(w * dt / 2))
    k_c = w / C0
    phase = np.angle(ratio)        # monitors sum e^{+i w t}: a wave delayed by D dz picks up e^{+i k D dz}
    print("phase", phase, "Yee", k_yee * D * dz, "continuum", k_c * D * dz, "|ratio|", np.abs(ratio), flush=True)
    assert np.all(np.abs(phase - k_yee * D * dz) < 1e-5)
    assert np.abs(phase[0] - k_c[0] * D * dz) > 4e-3
    assert np.all(np.abs(np.abs(ratio) - 1) < 1e-5)
