"""fdtd_run end to end on the device: an Input.txt in the reference format
in, SourceFile.txt, ScatAbs.<t>.txt and the snapshot files out
(main.cu:2629-2644, 2948-2990, 3042; Spectral.cu:1153-1171; snapshot.cu).
The spectra files are checked against the library run in this process on
the same parameters (columns 5-16 are fdtd_flux's face fluxes, 17-19 the
source spectrum; %e keeps 7 digits), and the normalised totals against the
reference formulas; the source file against the pulse from fdtd_get_info.
Then the files against the FP64 oracle run on the same Input.txt values:
ScatAbs columns 2-16, the |E| and |P_d| planes (YZ and XZ), the density
planes, the ey column and the 1D line of the aux_sim files."""
import os
import subprocess

import numpy as np
import pytest

import fdtd_amd
from cli_input import write_input

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fdtd-cuda_amd", "fdtd_run")
C0 = 2.99792458e8


def test_fdtd_run_matches_library(tmp_path):
    vals = write_input(tmp_path / "Input.txt", Tend=200, NCELLX=40, NCELLY=40, NCELLZ=40, NcpmlX=6, NcpmlY=6,
                       NcpmlZ=6, nano_sphere_radius=8e-9, NUM_freq=8, BandWidth=2.5e16, lam_0=5e-7,
                       Snap_in=3, t_skip=50, y_skip=2, z_skip=3)
    print("running fdtd_run on 53^3, Tend 200", flush=True)
    r = subprocess.run([EXE, "Input.txt"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    dumps = sorted(int(f.split(".")[1]) for f in os.listdir(tmp_path) if f.startswith("ScatAbs."))
    assert dumps == list(range(0, 201, 20))

    p = fdtd_amd.default_params()
    for k, v in vals.items():
        if hasattr(p, k):
            setattr(p, k, v)
    s = fdtd_amd.Simulation(p)
    info = s.get_info()
    s.step(201)
    fl = s.flux()
    raw = np.asarray(fl["raw"]).reshape(12, -1)
    freq, einc = s.spectrum_setup()
    out = np.loadtxt(tmp_path / "ScatAbs.200.txt")
    assert out.shape == (8, 19)
    np.testing.assert_allclose(out[:, 4:16], raw.T, rtol=2e-6, atol=1e-30)
    sca = -raw[0] + raw[1] - raw[2] + raw[3] - raw[4] + raw[5]
    ab = -raw[6] + raw[7] - raw[8] + raw[9] - raw[10] + raw[11]
    area = np.pi * vals["nano_sphere_radius"] ** 2
    np.testing.assert_allclose(out[:, 1], sca / area, rtol=2e-6)
    np.testing.assert_allclose(out[:, 2], -ab / area, rtol=2e-6)
    np.testing.assert_allclose(out[:, 3], sca / area - ab / area, rtol=2e-6, atol=1e-9 * np.abs(sca / area).max())
    np.testing.assert_allclose(out[:, 0], C0 / freq, rtol=2e-6)
    tol = 1e-6 * np.abs(einc).max()
    np.testing.assert_allclose(out[:, 16], np.abs(einc), rtol=2e-6, atol=tol)
    np.testing.assert_allclose(out[:, 17], einc.real, rtol=2e-6, atol=tol)
    np.testing.assert_allclose(out[:, 18], einc.imag, rtol=2e-6, atol=tol)

    # Snap_in 3 after steps 0, 50, ..., 200: the 1D files (count 0, 2, ..., 8), then |E| on the
    # plane i = NX/2, rows j::2, columns k::3 (count 1, 3, ..., 9)
    snaps = sorted(os.listdir(tmp_path / "videofiles"))
    assert snaps == sorted("YZ_sim.%d.txt" % n for n in range(1, 10, 2))
    assert sorted(f for f in os.listdir(tmp_path) if f.startswith("aux_sim.")) == sorted(
        ["aux_sim.%d.txt" % n for n in range(0, 10, 2)] + ["aux_sim.%d_2.txt" % n for n in range(0, 10, 2)])
    shot = np.loadtxt(tmp_path / "videofiles" / "YZ_sim.9.txt")
    e = [s.get_field(c)[info.NX // 2].astype(np.float64) for c in range(3)]
    mag = np.sqrt(e[0] ** 2 + e[1] ** 2 + e[2] ** 2)[::2, ::3]
    assert shot.shape == mag.shape and mag.max() > 0
    np.testing.assert_allclose(shot, mag, rtol=2e-6, atol=1e-6 * mag.max())

    src = np.loadtxt(tmp_path / "SourceFile.txt")
    assert src.shape == (200, 2)
    t = np.arange(200)
    b = ((t - info.delay) / info.width).astype(np.float32)
    pulse = vals["MAX_AMP"] * np.exp(-(b * b).astype(np.float64) / 2) * np.sin(2 * np.pi * info.f_0 * t * info.dt)
    np.testing.assert_allclose(src[:, 0], t * info.dt, rtol=2e-6)
    np.testing.assert_allclose(src[:, 1], pulse, rtol=2e-6, atol=1e-6 * np.abs(pulse).max())


def test_get_slice_matches_field():
    """fdtd_get_slice on each axis against the full-field copy."""
    import oracle as O
    from conftest import small_params
    p = small_params(O, steps=120)
    p.NCELLY = 26
    s = fdtd_amd.Simulation(p)
    s.step(120)
    for comp in (0, 4):
        f = s.get_field(comp)
        assert np.abs(f).max() > 0
        for axis, idx in ((0, 17), (1, 3), (2, 20)):
            want = np.take(f, idx, axis=axis)
            assert np.array_equal(s.get_slice(comp, axis, idx), want), (comp, axis)
    with pytest.raises(RuntimeError):
        s.get_slice(0, 2, s.info.NZ)


def test_fdtd_run_static_field_writes_no_spectra(tmp_path):
    """StaticField = 1 (ramped static source, no DFT): the reference writes
    no ScatAbs files (main.cu:3042); SourceFile.txt is still written."""
    write_input(tmp_path / "Input.txt", Tend=60, NCELLX=40, NCELLY=40, NCELLZ=40, NcpmlX=6, NcpmlY=6, NcpmlZ=6,
                nano_sphere_radius=8e-9, NUM_freq=4, StaticField=1)
    r = subprocess.run([EXE, "Input.txt", "--dir", "."], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert not [f for f in os.listdir(tmp_path) if f.startswith("ScatAbs.")]
    assert np.loadtxt(tmp_path / "SourceFile.txt").shape == (60, 2)


def test_density_slice_of_the_hydrodynamic_model():
    """fdtd_get_slice(6): N_EQ (n~x + n~y + n~z) / 3 on a plane; zero outside
    the metal box, nonzero inside once the pulse has driven the sphere, and
    the planes of two axes agree on the line where they cross.  (No FP64
    reference holds the density field; the values are pinned only by these
    properties and by the hydrodynamic field parity tests.)"""
    import oracle as O
    from conftest import small_params
    p = small_params(O, steps=300)
    p.NONLOCAL = 3
    p.hydro_beta = 1.27e6
    s = fdtd_amd.Simulation(p)
    s.step(300)
    info = s.get_info()
    c = info.NX // 2
    a = s.get_slice(6, 0, c)          # [j, k] at i = c
    b = s.get_slice(6, 1, c)          # [i, k] at j = c
    lo, hi = list(info.box_lo), list(info.box_hi)
    assert np.abs(a).max() > 0
    inside = np.zeros_like(a, dtype=bool)
    inside[lo[1]:hi[1], lo[2]:hi[2]] = True
    assert np.all(a[~inside] == 0)
    assert np.array_equal(a[c], b[c])


def oracle_params(vals):
    import oracle as O
    p = O.default_params()
    for k, v in vals.items():
        if hasattr(p, k):
            setattr(p, k, v)
    return p


def rel_to_max(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


def test_fdtd_run_spectra_and_snapshots_against_the_oracle(tmp_path):
    """Snap_in 3 with the XZ planes: ScatAbs.200.txt columns 2-16 (totals and
    the 12 face fluxes) within 1e-4 of the oracle's flux at their largest
    value (the scattering column at the largest extinction); the |E| planes YZ (i = NX/2) and XZ (j = NY/2) and the column
    ey(2, 2, k) within 2e-4 of the largest |E|; the 1D line to 1e-6 (FP64 on
    both sides, printed with 7 digits)."""
    import oracle as O
    vals = write_input(tmp_path / "Input.txt", Tend=200, NCELLX=40, NCELLY=40, NCELLZ=40, NcpmlX=6, NcpmlY=6,
                       NcpmlZ=6, nano_sphere_radius=8e-9, NUM_freq=8, BandWidth=2.5e16, lam_0=5e-7,
                       Snap_in=3, t_skip=100, x_skip=2, y_skip=3, z_skip=2)
    r = subprocess.run([EXE, "Input.txt", "--xz"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    o = O.Oracle(oracle_params(vals))
    o.step(201)
    info = o.get_info()
    fl = o.flux()
    out = np.loadtxt(tmp_path / "ScatAbs.200.txt")
    # scattering by an 8 nm sphere is two orders below its absorption: the
    # totals are judged against the largest extinction
    ext_max = np.abs(fl["ext"]).max()
    errs = {"sca": float(np.abs(out[:, 1] - fl["sca"]).max() / ext_max), "abs": rel_to_max(out[:, 2], fl["abs"]),
            "ext": rel_to_max(out[:, 3], fl["ext"]), "faces": rel_to_max(out[:, 4:16], fl["raw"].T)}
    e = [o.get_field(c) for c in range(3)]
    emax = max(np.abs(f).max() for f in e)
    mag = np.sqrt(e[0] ** 2 + e[1] ** 2 + e[2] ** 2)
    # t = 200 is the third snapshot block: aux_sim.4, then the planes numbered 5
    yz = np.loadtxt(tmp_path / "videofiles" / "YZ_sim.5.txt")
    xz = np.loadtxt(tmp_path / "videofiles" / "XZ_sim.5.txt")
    want_yz = mag[info.NX // 2, ::3, ::2]
    want_xz = mag[::2, info.NY // 2, ::2]
    assert yz.shape == want_yz.shape and xz.shape == want_xz.shape
    errs["YZ |E|"] = float(np.abs(yz - want_yz).max() / emax)
    errs["XZ |E|"] = float(np.abs(xz - want_xz).max() / emax)
    col = np.loadtxt(tmp_path / "aux_sim.4.txt")
    errs["ey(2,2,k)"] = float(np.abs(col - e[1][2, 2, ::2]).max() / emax)
    line = np.loadtxt(tmp_path / "aux_sim.4_2.txt")
    le, _ = o.get_line()
    errs["e_inc"] = rel_to_max(line, le[::2])
    print("fdtd_run vs oracle:", {k: "%.1e" % v for k, v in errs.items()}, flush=True)
    assert max(errs[k] for k in ("sca", "abs", "ext", "faces")) < 1e-4
    assert max(errs[k] for k in ("YZ |E|", "XZ |E|", "ey(2,2,k)")) < 2e-4
    assert errs["e_inc"] < 1e-6


@pytest.mark.parametrize("hydro", [0, 1])
def test_fdtd_run_drude_and_density_planes_against_the_oracle(tmp_path, hydro):
    """Snap_in 2 (main.cu:2954-2971): |P_d| of the Drude pole (the
    polarisation; the electron velocity with the hydrodynamic model) on the YZ
    and XZ planes, and the density planes (NDx + NDy + NDz) / 3 (zero without
    the hydrodynamic model), after the last step, within 2e-4 of the oracle's
    largest value."""
    import oracle as O
    vals = write_input(tmp_path / "Input.txt", Tend=240, NCELLX=40, NCELLY=40, NCELLZ=40, NcpmlX=6, NcpmlY=6,
                       NcpmlZ=6, nano_sphere_radius=8e-9, NUM_freq=4, BandWidth=2.5e16, lam_0=5e-7,
                       NONLOCAL=3 if hydro else 0, Snap_in=2, t_skip=240, x_skip=1, y_skip=2, z_skip=1)
    r = subprocess.run([EXE, "Input.txt", "--xz"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    o = O.Oracle(oracle_params(vals))   # hydro_beta is not an Input.txt token: both sides take the default
    assert o.get_info().hydro == hydro
    o.step(241)
    info = o.get_info()
    pd = [o.get_state(2, c) for c in range(3)]
    pmag = np.sqrt(pd[0] ** 2 + pd[1] ** 2 + pd[2] ** 2)
    nd = (o.get_state(0, 0) + o.get_state(0, 1) + o.get_state(0, 2)) / 3.0
    # t = 0 wrote planes 0 and 1_N, t = 240 planes 2 and 3_N
    vd = tmp_path / "videofiles"
    got = {"YZ |P_d|": (np.loadtxt(vd / "YZ_sim.2.txt"), pmag[info.NX // 2, ::2, :]),
           "XZ |P_d|": (np.loadtxt(vd / "XZ_sim.2.txt"), pmag[:, info.NY // 2, :]),
           "YZ N": (np.loadtxt(vd / "YZ_sim.3_N.txt"), nd[info.NX // 2, ::2, :]),
           "XZ N": (np.loadtxt(vd / "XZ_sim.3_N.txt"), nd[:, info.NY // 2, :])}
    errs = {}
    for k, (a, b) in got.items():
        assert a.shape == b.shape, k
        scale = np.abs(pmag).max() if "P_d" in k else max(np.abs(nd).max(), 1e-300)
        errs[k] = float(np.abs(a - b).max() / scale)
    print("Snap_in 2, hydro %d vs oracle:" % hydro, {k: "%.1e" % v for k, v in errs.items()},
          "max |P_d| %.3e, max |N| %.3e" % (np.abs(pmag).max(), np.abs(nd).max()), flush=True)
    assert np.abs(pmag).max() > 0
    assert errs["YZ |P_d|"] < 2e-4 and errs["XZ |P_d|"] < 2e-4
    if hydro:
        assert np.abs(nd).max() > 0
        assert errs["YZ N"] < 2e-4 and errs["XZ N"] < 2e-4
    else:
        assert np.all(got["YZ N"][0] == 0) and np.all(got["XZ N"][0] == 0)
