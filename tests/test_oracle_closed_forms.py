"""Pins the FP64 oracle against closed forms and invariants (CPU only).

The reference ships no tests or golden data, so each oracle function is pinned
by mathematics chosen so a plausible transcription mistake (a swapped slab, a
dropped term, a wrong sign or offset) fails: CPML profiles against the
graded-profile formulas, the discrete ADE against the Drude + critical-point
permittivity, the 1D line against the vacuum speed of light and Mur
absorption, TFSF leakage in vacuum, and DFT sums against geometric series.
"""
import numpy as np
import pytest

import oracle as O
from conftest import small_params

EP0, C0, Z0, MU0 = 8.85418782e-12, 2.99792458e8, 376.730313462, 1.25663706e-06


def make(**kw):
    p = small_params(O, **kw)
    return O.Oracle(p), p


# ---------------------------------------------------------------- CPML profiles
def test_cpml_profiles_closed_form():
    o, p = make()
    dt = o.coeffs()["dt"]
    d, Nc, m = p.dx, p.NcpmlX, p.cpml_exp
    smax = p.sigma_factor * (m + 1) / (Z0 * d)

    def bc(x, xa):
        S = x ** m * smax
        K = 1 + (p.kappa_max - 1) * x ** m
        A = p.alpha_max * xa
        b = np.exp(-(S / K + A) * dt / EP0)
        c = np.where(K * A + S == 0, 0.0, S * (b - 1) / (K * (S + K * A)))
        return b, c, K

    i = np.arange(Nc + 1, dtype=float)
    h = np.arange(Nc, dtype=float)
    for axis in range(3):
        b, c, K = bc((Nc - i) / Nc, i / Nc)
        np.testing.assert_allclose(o.cpml(axis, "be_N"), b, rtol=1e-14)
        np.testing.assert_allclose(o.cpml(axis, "ce_N"), c, rtol=1e-14, atol=1e-300)
        b, c, _ = bc(i / Nc, (Nc - i) / Nc)
        np.testing.assert_allclose(o.cpml(axis, "be_F"), b, rtol=1e-14)
        np.testing.assert_allclose(o.cpml(axis, "ce_F"), c, rtol=1e-14, atol=1e-300)
        b, c, Kh = bc((Nc - h - 0.5) / Nc, (h + 0.5) / Nc)
        np.testing.assert_allclose(o.cpml(axis, "bh_N"), b, rtol=1e-14)
        np.testing.assert_allclose(o.cpml(axis, "ch_N"), c, rtol=1e-14)
        ked, khd = o.cpml(axis, "ked"), o.cpml(axis, "khd")
        # kappa peaks at the walls (kappa_max on the outermost E cell), 1 inside
        assert ked[0] == pytest.approx(p.kappa_max * d) and ked[-1] == pytest.approx(p.kappa_max * d)
        np.testing.assert_allclose(ked[Nc:len(ked) - Nc - 1], d)
        np.testing.assert_allclose(khd[:Nc], Kh * d, rtol=1e-14)
        # no conductivity at the inner interface: c == 0 there
        assert o.cpml(axis, "ce_N")[Nc] == 0.0 and o.cpml(axis, "ce_F")[0] == 0.0


# ---------------------------------------------------------------- dispersive model
def drude_cp_eps(w, conj_convention=True):
    """Gold Drude + 2 CP permittivity (Vial-Laroche form, exp(-i w t)); conjugated for exp(+i w t)."""
    eps_inf, wD, gD = 1.1431, 1.3202e16, 1.0805e14
    A = [0.26698, 3.0834]; phi = [-1.2371, -1.0968]; Om = [3.8711e15, 4.1684e15]; Ga = [4.4642e14, 2.3555e15]
    e = eps_inf - wD ** 2 / (w ** 2 + 1j * gD * w)
    for a, f, o, g in zip(A, phi, Om, Ga):
        e = e + a * o * (np.exp(1j * f) / (o - w - 1j * g) + np.exp(-1j * f) / (o + w + 1j * g))
    return np.conj(e) if conj_convention else e


def discrete_eps(c, w, hydro=False):
    """Permittivity of the reference's discrete E/ADE update for E, P ~ z^n, z = exp(i w dt)."""
    dt = c["dt"]
    z = np.exp(1j * w * dt)
    chi = 0
    poles = [(c["d1"], c["d2"], c["d3"], c["d4"], c["d5"])] + [
        tuple(c[f"{n}_{i}"] for n in ("c1", "c2", "c3", "c4", "c5")) for i in range(2)]
    for k1, k2, k3, k4, k5 in poles:
        chi = chi + (k3 * z + k4 + k5 / z) / (z - k1 - k2 / z)
    return c["eps_inf"] + chi / EP0


def test_ade_reproduces_drude_cp_permittivity():
    o, _ = make()
    c = o.coeffs()
    lam = np.linspace(400e-9, 1000e-9, 13)
    w = 2 * np.pi * C0 / lam
    ed = discrete_eps(c, w)
    ec = drude_cp_eps(w)
    wdt = w * c["dt"]
    # second-order accurate: relative error bounded by a small multiple of (w dt)^2
    err = np.abs(ed - ec) / np.abs(ec)
    assert np.all(err < 5 * wdt ** 2 + 1e-9), (err, wdt)
    # gold near 600 nm: strongly negative real part, modest loss
    e600 = drude_cp_eps(2 * np.pi * C0 / 600e-9)
    assert -11 < e600.real < -7 and 0.5 < abs(e600.imag) < 3


def test_ade_update_coefficients_consistent():
    """C_E, C_E_1, C_E_2 must be the sums implied by eps0*eps_inf and the pole drives."""
    o, _ = make()
    c = o.coeffs()
    s3 = c["d3"] + c["c3_0"] + c["c3_1"]
    s4 = c["d4"] + c["c4_0"] + c["c4_1"]
    s5 = c["d5"] + c["c5_0"] + c["c5_1"]
    assert c["C_E"] == pytest.approx(EP0 * c["eps_inf"] + s3, rel=1e-15)
    assert c["C_E_1"] == pytest.approx(EP0 * c["eps_inf"] - s4, rel=1e-15)
    assert c["C_E_2"] == pytest.approx(s5, rel=1e-15)
    assert c["Che"] == pytest.approx(Z0 * c["dt"] / MU0, rel=1e-15)
    assert c["Ceh"] == pytest.approx(c["dt"] / (Z0 * EP0), rel=1e-15)


# ---------------------------------------------------------------- 1D incident line
def test_line_pulse_speed_and_mur():
    o, p = make(sphere=0, steps=600)
    info = o.info
    m0, L = info.m0, info.inc_length
    dt, dx = info.dt, p.dx
    courant = C0 * dt / dx
    src = m0 - p.source_offset
    peak_t = None
    best = 0
    for t in range(600):
        o.step(1)
        e, _ = o.get_line()
        if abs(e[m0]) > best:
            best, peak_t = abs(e[m0]), t
    # the Gaussian-modulated sine envelope peaks at delay and travels p.source_offset cells to m0;
    # the carrier shifts the largest sample by up to half a period
    period = 1.0 / (info.f_0 * dt)
    expect = info.delay + p.source_offset / courant
    assert abs(peak_t - expect) < period / 2 + 3, (peak_t, expect)
    assert src > 0 and m0 < L


def test_line_mur_absorbs():
    o, p = make(sphere=0, steps=1400)
    peak = 0.0
    for t in range(1400):
        o.step(1)
        if t % 10 == 0:
            peak = max(peak, np.max(np.abs(o.get_line()[0])))
    e, _ = o.get_line()
    # after both wave packets have left, first-order Mur leaves a small residue
    assert np.max(np.abs(e)) < 2e-2 * peak


# ---------------------------------------------------------------- TFSF
def test_tfsf_leakage_in_vacuum():
    """No scatterer: the scattered-field region outside the TFSF box stays near zero."""
    for phi in (0.0, 90.0):
        o, p = make(sphere=0, steps=330, phi=phi)
        o.step(330)
        info = o.info
        t0 = np.array(info.tfsf[:])
        N = np.array([info.NX, info.NY, info.NZ])
        t1 = N - t0
        inside = (slice(t0[0] + 1, t1[0] - 1), slice(t0[1] + 1, t1[1] - 1), slice(t0[2] + 1, t1[2] - 1))
        total = max(np.max(np.abs(o.get_field(c)[inside])) for c in range(6))
        assert total > 1e-3
        outside = np.ones(o.shape, bool)
        outside[t0[0] - 2:t1[0] + 2, t0[1] - 2:t1[1] + 2, t0[2] - 2:t1[2] + 2] = False
        leak = max(np.max(np.abs(o.get_field(c)[outside])) for c in range(6))
        assert leak < 1e-3 * total, (phi, leak, total)


def test_polarisation_and_direction():
    """phi = 0 drives E along -y (tfsf2.cu:11-13); phi = 90 drives E along +x."""
    o, _ = make(sphere=0, steps=320)
    o.step(320)
    ex, ey = np.max(np.abs(o.get_field(0))), np.max(np.abs(o.get_field(1)))
    assert ey > 1e3 * ex
    o2, _ = make(sphere=0, steps=320, phi=90.0)
    o2.step(320)
    ex2, ey2 = np.max(np.abs(o2.get_field(0))), np.max(np.abs(o2.get_field(1)))
    assert ex2 > 1e3 * ey2


# ---------------------------------------------------------------- DFT
def test_dft_geometric_sums():
    """Constant fields: each face accumulates sum cos/sin(2 pi f tau) over sampled steps."""
    o, p = make(sphere=0, steps=50)
    shape = o.shape
    vals = [1.0, 2.0, 3.0, 4.0, 5.0, 6.0]
    nsteps = 37
    for t in range(nsteps):
        for c in range(6):
            o.set_field(c, np.full(shape, vals[c]))
        o.end_step()
    freq, _ = o.spectrum_setup()
    dt = o.info.dt
    ts = np.array([t for t in range(nsteps) if t % p.dft_stride == 0])
    se = np.exp(1j * 2 * np.pi * np.outer(freq, ts + 1.0) * dt).sum(1)
    sh = np.exp(1j * 2 * np.pi * np.outer(freq, ts + 0.5) * dt).sum(1)
    for face in range(12):
        axis = (face // 2) % 3
        a, b = (axis + 1) % 3, (axis + 2) % 3
        expect = [vals[a] * se, vals[b] * se, vals[3 + a] * sh, vals[3 + b] * sh]
        for comp in range(4):
            d = o.get_dft(face, comp)
            np.testing.assert_allclose(d, np.broadcast_to(expect[comp][:, None, None], d.shape), rtol=1e-12)


def test_flux_vacuum_is_small():
    o, p = make(sphere=0, steps=650, nfreq=6)
    o.step(650)
    f = o.flux()
    raw = f["raw"]
    # net flux through each closed box is tiny compared with one face's incident flux
    ref = np.max(np.abs(raw[6 + 4:6 + 6]))
    assert ref > 0
    assert np.max(np.abs(f["sca"] * np.pi * p.nano_sphere_radius ** 2)) < 1e-2 * ref
    assert np.max(np.abs(f["abs"] * np.pi * p.nano_sphere_radius ** 2)) < 1e-2 * ref


def test_source_spectrum():
    o, p = make(steps=300)
    freq, einc = o.spectrum_setup()
    info = o.info
    t = np.arange(p.Tend)
    b = ((t - info.delay) / info.width).astype(np.float32)
    pulse = p.MAX_AMP * np.exp(-(b * b).astype(np.float64) / 2) * np.sin(2 * np.pi * info.f_0 * t * info.dt)
    expect = (pulse[None, :] * np.exp(1j * 2 * np.pi * np.outer(freq, t) * info.dt)).sum(1)
    np.testing.assert_allclose(einc, expect, rtol=1e-9)
    lam = np.linspace(p.lam_min, p.lam_max, p.NUM_freq)
    np.testing.assert_allclose(freq, C0 / lam, rtol=1e-14)


def test_hydro_local_limit_is_drude():
    """NONLOCAL = 3 with the pressure term off (beta = 0) and linear charge:
    the discrete velocity + current update must reproduce the Drude pole with
    wp^2 = e^2 N_EQ / (eps0 m_e), to first order in w dt (the velocity is
    driven by E^n and E^{n-1})."""
    E0, ME = 1.60217662e-19, 9.10938356e-31
    o, p = make()
    p.NONLOCAL = 3
    p.hydro_beta = 0.0
    oh = O.Oracle(p)
    c = oh.coeffs()
    assert c["dNL"] == 0.0
    dt, NEQ = c["dt"], c["N_EQ"]
    lam = np.linspace(400e-9, 1000e-9, 13)
    w = 2 * np.pi * C0 / lam
    z = np.exp(1j * w * dt)
    v_over_e = (c["d3"] + c["d4"] / z) / (z - c["d1"] - c["d2"] / z)
    chi_h = -dt * E0 * NEQ * z * v_over_e / (EP0 * (z - 1))
    chi_cp = 0
    for i in range(2):
        k1, k2, k3, k4, k5 = (c[f"{n}_{i}"] for n in ("c1", "c2", "c3", "c4", "c5"))
        chi_cp = chi_cp + (k3 * z + k4 + k5 / z) / (z - k1 - k2 / z) / EP0
    eps_h = c["eps_inf"] + chi_cp + chi_h
    wp2 = E0 ** 2 * NEQ / (EP0 * ME)
    assert np.sqrt(wp2) == pytest.approx(1.3202e16, rel=2e-4)
    # continuum Drude with this wp and the gold damping, plus the same CP poles
    ref = drude_cp_eps(w) + (1.3202e16 ** 2 - wp2) / (w ** 2 - 1j * 1.0805e14 * w)
    err = np.abs(eps_h - ref) / np.abs(ref)
    assert np.all(err < 5 * w * dt), (err, w * dt)


def test_point_objective_terms_are_the_dft_of_the_recorded_field():
    """The oracle's point-objective sums equal a numpy DFT, with phase
    (t+1) dt, of E at r0 read back after every step."""
    p = small_params(O, n=18, ncpml=4, steps=60, sphere=0, phi=90.0)
    p.BandWidth = 8e16
    o = O.Oracle(p)
    info = o.get_info()
    r0 = (info.NX // 2, info.NY // 2 + 1, info.NZ // 2 - 2)
    freqs = np.array([3.5e14, 4.7e14, 6.1e14])
    o.set_objective(r0, freqs)
    dt = info.dt
    rec = []
    for t in range(60):
        o.step(1)
        rec.append([o.get_field(c)[r0] for c in range(3)])
    rec = np.array(rec)
    assert np.abs(rec).max() > 0
    t = np.arange(60)
    w = np.exp(1j * 2 * np.pi * freqs[:, None] * (t[None, :] + 1.0) * dt)
    ref = w @ rec
    got = o.objective_terms()
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    assert o.objective() == pytest.approx(float(np.sum(np.abs(ref) ** 2)), rel=1e-12)


# ---------------------------------------------------------------- Lorentz-Drude (material 7)
def make_with(**over):
    p = small_params(O)
    for k, v in over.items():
        setattr(p, k, v)
    return O.Oracle(p), p


def lorentz_drude_eps(w):
    """Lorentz-Drude permittivity of material 7 (setup.cu:1256-1275), exp(-i w t):
    eps_inf - wD^2/(w^2 + i gD w) + sum_L deps_L wL^2/(wL^2 - w^2 - 2 i dL w),
    parameters in eV converted with 2 pi / h."""
    h = 4.13566553853599E-15
    c = 2 * np.pi / h
    e = 3.559 - (c * 8.812) ** 2 / (w ** 2 + 1j * c * 0.0752 * w)
    for de, dl, wl in ((2.912, 1.541, 4.693), (1.272, 0.525, 3.112)):
        e = e + de * (c * wl) ** 2 / ((c * wl) ** 2 - w ** 2 - 2j * c * dl * w)
    return e


def test_lorentz_drude_update_reproduces_its_permittivity():
    """Material 7: the z-transform of the five-coefficient pole recursions
    with the E update built from their sums (C_E, C_E_1, C_E_2) equals the
    continuous Lorentz-Drude permittivity to O((w dt)^2); also the sums."""
    o, _ = make_with(material=7)
    c = o.coeffs()
    lam = np.linspace(400e-9, 1000e-9, 13)
    w = 2 * np.pi * C0 / lam
    ed = discrete_eps(c, w)
    ec = np.conj(lorentz_drude_eps(w))
    err = np.abs(ed - ec) / np.abs(ec)
    wdt = w * c["dt"]
    assert np.all(err < 5 * wdt ** 2 + 1e-9), (err, wdt)
    assert c["eps_inf"] == 3.559
    S3 = c["d3"] + c["c3_0"] + c["c3_1"]
    assert c["C_E"] == pytest.approx(EP0 * 3.559 + S3, rel=1e-14)
    # the Drude pole keeps k1 + k2 = 1 (no restoring force)
    assert c["d1"] + c["d2"] == pytest.approx(1.0, abs=1e-15)


def test_nonlocal_pole_with_zero_velocity_is_the_local_model():
    """NONLOCAL = 1 with beta = 0 must give the local run's fields bit for bit."""
    out = []
    for nl in (0, 1):
        o, p = make_with(NONLOCAL=nl, hydro_beta=0.0, Tend=200)
        o.step(200)
        out.append([o.get_field(k) for k in range(6)])
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("material", [3, 7])
def test_nonlocal_pole_stays_bounded(material):
    """A sign mistake in the grad-div term turns the longitudinal modes
    unstable (w^2 = wp^2 - beta^2 k^2 < 0 on the grid's shortest waves): with
    a large beta the run must stay finite and decay after the pulse."""
    o, p = make_with(NONLOCAL=1, hydro_beta=8e6, material=material, Tend=900)
    peaks = []
    for _ in range(3):
        o.step(300)
        peaks.append(max(np.abs(o.get_field(k)).max() for k in range(3)))
    assert np.all(np.isfinite(peaks)), peaks
    assert peaks[2] < 2 * max(peaks[:2]), peaks
