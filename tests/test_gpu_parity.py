"""HIP solver against the FP64 oracle on small grids, through the C ABI.

Tolerances: fields are FP32 on the device against FP64 in the oracle.  The
error of a component is its L-inf difference normalised by the largest value
of any E (or H) component, so cross-polarised components that are nearly zero
are judged against the field that drives them.  After 400 steps (the pulse has
crossed the TFSF box) that error sits near 1e-6..1e-5; the bar is 2e-4.
"""
import numpy as np
import pytest

import fdtd_amd
import oracle as O
from conftest import small_params

pytestmark = pytest.mark.gpu

TOL = 2e-4


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def field_errors(s, o):
    fo = [o.get_field(c) for c in range(6)]
    fs = [s.get_field(c) for c in range(6)]
    me = max(np.abs(f).max() for f in fo[:3])
    mh = max(np.abs(f).max() for f in fo[3:])
    return [float(np.abs(fs[c] - fo[c]).max() / (me if c < 3 else mh)) for c in range(6)]


def run_pair(p, steps, fused):
    o = O.Oracle(p)
    s = fdtd_amd.Simulation(p)
    o.step(steps)
    if fused:
        s.step(steps)
    else:
        s.split_step(steps)
    return o, s


@pytest.mark.parametrize("fused", [False, True])
def test_fields_gold_sphere(fused):
    p = small_params(O, steps=400)
    o, s = run_pair(p, 400, fused)
    assert s.get_info().n_dispersive[0] == o.info.n_dispersive[0] > 0
    errs = field_errors(s, o)
    print(f"gold sphere fused={fused} errors={['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) < TOL, errs


@pytest.mark.parametrize("fused", [False, True])
def test_fields_vacuum_x_polarised(fused):
    p = small_params(O, sphere=0, phi=90.0, steps=400)
    o, s = run_pair(p, 400, fused)
    errs = field_errors(s, o)
    print(f"vacuum fused={fused} errors={['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) < TOL, errs


def test_line_matches_oracle():
    p = small_params(O, steps=100)
    o, s = run_pair(p, 100, True)
    eo, ho = o.get_line()
    es, hs = s.get_line()
    assert rel(es, eo) < 1e-12
    assert rel(hs, ho) < 1e-12


@pytest.mark.parametrize("fused", [False, True])
def test_dft_and_flux(fused):
    """Running DFTs of the tangential fields on the 12 monitor faces and the
    flux spectra.  A DFT error is normalised by the largest E (or H) DFT value
    over all faces, as field errors are; spectra by their own maximum."""
    p = small_params(O, steps=450, block=8)
    o, s = run_pair(p, 449, fused)   # 113 samples: ends mid-block, exercises the partial contraction
    da = {(f, c): s.get_dft(f, c) for f in range(12) for c in range(4)}
    db = {(f, c): o.get_dft(f, c) for f in range(12) for c in range(4)}
    me = max(np.abs(db[f, c]).max() for f in range(12) for c in (0, 1))
    mh = max(np.abs(db[f, c]).max() for f in range(12) for c in (2, 3))
    worst, own = 0.0, 0.0
    for (f, c), a in da.items():
        err = float(np.abs(a - db[f, c]).max())
        worst = max(worst, err / (me if c < 2 else mh))
        own = max(own, rel(a, db[f, c]))
    print(f"dft worst rel={worst:.3e} (per face-component normalisation {own:.3e})", flush=True)
    assert worst < 5e-5
    fo, fs = o.flux(), s.flux()
    for k in ("raw", "sca", "abs", "ext"):
        err = rel(fs[k], fo[k])
        print(f"flux {k} rel={err:.3e}", flush=True)
        assert err < 1e-4, (k, err)


def group_errors(fs, fo):
    """L-inf differences normalised by the largest E (or H) value, as field_errors."""
    me = max(np.abs(f).max() for f in fo[:3])
    mh = max(np.abs(f).max() for f in fo[3:])
    return [float(np.abs(fs[c] - fo[c]).max() / (me if c < 3 else mh)) for c in range(6)]


def test_fused_equals_split():
    """The fused step (TFSF added after the update, linear in the fields)
    against the split path (TFSF between the H and E updates, reference
    order): the same fields up to FP32 rounding."""
    p = small_params(O, steps=400)
    a = fdtd_amd.Simulation(p)
    b = fdtd_amd.Simulation(p)
    a.step(400)
    b.split_step(400)
    errs = group_errors([a.get_field(c) for c in range(6)], [b.get_field(c) for c in range(6)])
    print(f"fused vs split errors={['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) < 1e-5


@pytest.mark.parametrize("ctj,rows,cchunk", [(4, 2, 5), (8, 3, 3), (16, 2, 9), (8, 2, 64)])
def test_tile_shapes_agree(ctj, rows, cchunk, monkeypatch):
    """Tile shapes of the row path (rows per thread, clean-box x chunk) and of
    the general kernel on the corner columns (FDTD_CTJ tile rows) give the
    split path's fields."""
    p = small_params(O, steps=350)
    ref = fdtd_amd.Simulation(p)
    ref.split_step(350)
    for k, v in (("FDTD_CTJ", ctj), ("FDTD_ROWS", rows), ("FDTD_CCHUNK", cchunk)):
        monkeypatch.setenv(k, str(v))
    s = fdtd_amd.Simulation(p)
    assert s.get_info().fused_path == 1
    s.step(350)
    errs = group_errors([s.get_field(c) for c in range(6)], [ref.get_field(c) for c in range(6)])
    print(f"tile ctj={ctj} rows={rows} cchunk={cchunk} errors={['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) < 1e-5


@pytest.mark.parametrize("tj,ichunk", [(4, 7), (8, 3), (16, 5), (8, 32)])
def test_general_kernel_tile_shapes_agree(tj, ichunk, monkeypatch):
    """Every tile of the general fused kernel over the whole grid
    (FDTD_PATH=general: TJ rows per block, x chunk) gives the split path's
    fields."""
    p = small_params(O, steps=350)
    ref = fdtd_amd.Simulation(p)
    ref.split_step(350)
    for k, v in (("FDTD_PATH", "general"), ("FDTD_TJ", tj), ("FDTD_ICHUNK", ichunk)):
        monkeypatch.setenv(k, str(v))
    s = fdtd_amd.Simulation(p)
    assert s.get_info().fused_path == 0
    s.step(350)
    errs = group_errors([s.get_field(c) for c in range(6)], [ref.get_field(c) for c in range(6)])
    print(f"general tj={tj} ichunk={ichunk} errors={['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) < 1e-5


@pytest.mark.parametrize("sphere", [1, 0])
def test_general_kernel_alone(sphere, monkeypatch):
    """FDTD_PATH=general: the general fused kernel over the whole grid with
    TFSF inside, against the oracle."""
    monkeypatch.setenv("FDTD_PATH", "general")
    p = small_params(O, steps=400, sphere=sphere, phi=0.0 if sphere else 90.0)
    o, s = run_pair(p, 400, True)
    errs = field_errors(s, o)
    print(f"general path sphere={sphere} errors={['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) < TOL, errs


def test_uneven_grid_and_cpml():
    """Unequal edges and CPML thicknesses (clean region bounds differ per axis)."""
    p = small_params(O, steps=360)
    p.NCELLX, p.NCELLY, p.NCELLZ = 22, 26, 30
    p.NcpmlX, p.NcpmlY, p.NcpmlZ = 5, 7, 6
    o, s = run_pair(p, 360, True)
    errs = field_errors(s, o)
    print(f"uneven errors={['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) < TOL, errs


@pytest.mark.parametrize("mode,tol", [(0, 1e-4), (2, 2e-6)])
def test_mfma_split_bf16_matches_fp64_contraction(mode, tol):
    """The split-BF16 MFMA contraction against the exact FP64 contraction on the
    same samples: checks the MFMA operand layout end to end and the split
    precision (two parts, hi*hi + hi*lo + lo*hi: ~2^-16 relative error per
    product; three parts, six products: ~2^-24, FP32 level)."""
    p = small_params(O, steps=450, block=16, nfreq=24)
    p.dft_exact = mode
    a = fdtd_amd.Simulation(p)
    p2 = small_params(O, steps=450, block=16, nfreq=24)
    p2.dft_exact = 1
    b = fdtd_amd.Simulation(p2)
    a.step(450)
    b.step(450)
    worst = 0.0
    for face in range(12):
        for comp in range(4):
            worst = max(worst, rel(a.get_dft(face, comp), b.get_dft(face, comp)))
    print(f"mfma ({2 if mode == 0 else 3} parts) vs fp64 contraction worst rel={worst:.3e}", flush=True)
    assert worst < tol
    fa, fb = a.flux(), b.flux()
    assert rel(fa["ext"], fb["ext"]) < tol


@pytest.mark.parametrize("fused", [False, True])
def test_hydrodynamic_sphere(fused):
    """NONLOCAL = 3: velocity/charge state on the metal box against the oracle."""
    p = small_params(O, steps=400)
    p.NONLOCAL = 3
    o, s = run_pair(p, 400, fused)
    assert s.get_info().hydro == 1
    errs = field_errors(s, o)
    print(f"hydro fused={fused} errors={['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) < TOL, errs


def test_point_monitor_gap_field_of_a_dimer():
    """|E(f)|^2 at the gap centre of a small gold dimer (the gap-field metric)
    against the oracle's FP64 sums; normalised by the largest value, 1e-4."""
    p = small_params(O, n=30, ncpml=6, sphere=0, steps=500, phi=90.0)
    s = fdtd_amd.Simulation(p)
    T = s.shape[0]
    c = T // 2
    # 3-cell spheres ((int)(6.5 nm / 2 nm)) centred 8 cells apart: cells c-1..c+1 of the gap are vacuum
    s.add_sphere((c - 4, c, c), 6.5e-9)
    s.add_sphere((c + 4, c, c), 6.5e-9)
    o = O.Oracle(p)
    o.set_materials(s.get_materials(1), s.get_materials(2), s.get_materials(3))
    r0 = (c, c, c)
    freqs = np.array([3.8e14, 4.6e14, 5.5e14, 6.7e14])
    s.set_point_monitor(r0, 100)
    s.set_point_monitor(r0, 600)   # a larger capacity replaces the first buffer
    o.set_objective(r0, freqs)
    s.step(500)
    o.step(500)
    e2, amp = s.point_spectrum(freqs)
    ref = o.objective_terms()
    e2o = np.sum(np.abs(ref) ** 2, axis=1)
    err = float(np.max(np.abs(e2 - e2o)) / np.max(e2o))
    aerr = float(np.max(np.abs(amp - ref)) / np.max(np.abs(ref)))
    print("gap |E|^2", e2, "oracle", e2o, "err %.2e amp err %.2e" % (err, aerr), flush=True)
    assert s.get_info().n_dispersive[0] > 0
    assert err < 1e-4 and aerr < 1e-4


@pytest.mark.parametrize("hydro,radius,offsets,schedule", [(0, 8e-9, (2, 2), 3), (0, 4e-9, (3, 1), 7),
                                                           (1, 8e-9, (2, 2), 3), (1, 4e-9, (3, 1), 7)])
def test_step_schedules_bit_identical(hydro, radius, offsets, schedule, monkeypatch):
    """fdtd_step with the side pieces on three streams (FDTD_SIDE_STREAMS=1:
    one) beside the clean box, with the TFSF fix-up and the dispersive box
    (schedule bits 0, 1) and the DFT sample (bit 2) following the clean box
    ahead of the side pieces, in one launch with the Drude metal, against
    them after the join (FDTD_EARLY=0) and on one stream (FDTD_CONCURRENT=0):
    the cells each kernel touches are disjoint, so every schedule gives the
    same bits, DFT spectra included; the default starts the next step's x
    and y slabs and corners at the join (schedule bit 3 where the cells allow
    it; the line then runs first on the corner stream), FDTD_EARLY_SIDE=1
    only the y slabs and corners, 0 none.  The sample joins only when its faces
    clear the TFSF planes (scattering offset 2 puts the far face on plane
    t1+1; 3 clears it) and the dispersive box with its margin (the 4 nm
    sphere's box clears absorption faces at offset 1)."""
    p = small_params(O, steps=300, block=8, radius=radius)
    p.sca_offset, p.abs_offset = offsets
    if hydro:
        p.NONLOCAL = 3
        p.hydro_beta = 1.27e6
    runs = {}
    for name, env in (("early", {}), ("late", {"FDTD_EARLY": "0"}), ("one side stream", {"FDTD_SIDE_STREAMS": "1"}),
                      ("no early side", {"FDTD_EARLY_SIDE": "0"}), ("early y slabs only", {"FDTD_EARLY_SIDE": "1"}),
                      ("serial", {"FDTD_CONCURRENT": "0"})):
        for k in ("FDTD_EARLY", "FDTD_CONCURRENT", "FDTD_SIDE_STREAMS", "FDTD_EARLY_SIDE"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        s = fdtd_amd.Simulation(p)
        info = s.get_info()
        assert info.hydro == hydro and info.n_dispersive[0] > 0
        s.step(300)
        runs[name] = (info.schedule, [s.get_field(c) for c in range(6)], [s.get_dft(f, 0) for f in range(12)])
    assert runs["early"][0] & 7 == schedule, runs["early"][0]
    assert runs["late"][0] == 0 and runs["serial"][0] == 0 and runs["one side stream"][0] & 7 == schedule
    assert runs["early"][0] & 8, runs["early"][0]
    assert runs["no early side"][0] == schedule, runs["no early side"][0]
    _, ref, ref_dft = runs["serial"]
    assert runs["early y slabs only"][0] & 8
    for name in ("early", "late", "one side stream", "no early side", "early y slabs only"):
        _, fields, dft = runs[name]
        for c in range(6):
            assert np.array_equal(fields[c], ref[c]), (name, c)
        for f in range(12):
            assert np.array_equal(dft[f], ref_dft[f]), (name, f)


@pytest.mark.parametrize("ny,rows,tiles,stiles", [(24, 2, 3, 1), (26, 2, 3, 1), (28, 2, 3, 1), (26, 3, 3, 1),
                                                (26, 2, 2, 2), (28, 3, 2, 2), (27, 2, 1, 2)])
def test_row_tiles_per_block_bit_identical(ny, rows, tiles, stiles, monkeypatch):
    """Two or three row tiles per block on the clean box (FDTD_TILES) and two
    on the side slabs (FDTD_STILES) against one: the same cells with the same
    arithmetic, so the same bits, also when the row-tile count is not a
    multiple of the tiles (idle tiles store nothing)."""
    p = small_params(O, steps=300)
    p.NCELLY = ny
    monkeypatch.setenv("FDTD_ROWS", str(rows))
    out = {}
    for t, st in ((1, 1), (tiles, stiles)):
        monkeypatch.setenv("FDTD_TILES", str(t))
        monkeypatch.setenv("FDTD_STILES", str(st))
        s = fdtd_amd.Simulation(p)
        s.step(300)
        out[t, st] = [s.get_field(c) for c in range(6)]
    assert np.abs(out[tiles, stiles][0]).max() > 0
    for c in range(6):
        assert np.array_equal(out[1, 1][c], out[tiles, stiles][c]), c


@pytest.mark.parametrize("phi", [0.0, 90.0])
def test_hydrodynamic_density_and_drude_pole_planes(phi):
    """The hydrodynamic state read back through fdtd_get_slice against the
    oracle: the density (NDx + NDy + NDz)/3 (component 6) and the Drude pole,
    here the electron velocity (components 7-9), on the planes through the
    sphere's centre along each axis, after the pulse has driven the sphere.
    The density barely touches the fields (it enters as 1 + n~), so the field
    tests do not see it; this one does (within 1e-4 of the largest value)."""
    p = small_params(O, steps=300)
    p.NONLOCAL = 3
    p.hydro_beta = 1.27e6
    p.inc_phi = phi
    o, s = run_pair(p, 300, True)
    c = [n // 2 for n in o.shape]
    nd = sum(o.get_state(0, q) for q in range(3)) / 3.0
    pd = [o.get_state(2, q) for q in range(3)]
    errs = {}
    for axis in range(3):
        want = np.take(nd, c[axis], axis=axis)
        errs["N axis %d" % axis] = float(np.abs(s.get_slice(6, axis, c[axis]) - want).max() / np.abs(nd).max())
        for q in range(3):
            w = np.take(pd[q], c[axis], axis=axis)
            errs["P%d axis %d" % (q, axis)] = float(np.abs(s.get_slice(7 + q, axis, c[axis]) - w).max() / np.abs(pd[q]).max())
    print("phi %g density / velocity planes:" % phi, {k: "%.1e" % v for k, v in errs.items()}, flush=True)
    assert np.abs(nd).max() > 0
    assert max(errs.values()) < 1e-4
