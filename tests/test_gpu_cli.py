"""fdtd_run end to end on the device: an Input.txt in the reference format
in, SourceFile.txt, ScatAbs.<t>.txt and the YZ |E| snapshots out
(main.cu:2629-2644, 3042; Spectral.cu:1153-1171; snapshot.cu:62-76).  The spectra files are checked against the library
run in this process on the same parameters (columns 5-16 are fdtd_flux's
face fluxes, 17-19 the source spectrum; %e keeps 7 digits), and the
normalised totals against the reference formulas; the source file against
the pulse from fdtd_get_info."""
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
                       Snap_in=2, t_skip=50, y_skip=2, z_skip=3)
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

    # snapshots after steps 0, 50, ..., 200: |E| on the plane i = NX/2, rows j::2, columns k::3
    snaps = sorted(os.listdir(tmp_path / "videofiles"))
    assert snaps == ["YZ_sim.%d.txt" % n for n in range(5)]
    shot = np.loadtxt(tmp_path / "videofiles" / "YZ_sim.4.txt")
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
