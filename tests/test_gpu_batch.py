"""Batched sweeps (fdtd_batch_*) through the C ABI.

* Every simulation of an uneven-size batch against the FP64 oracle (fields
  2e-4 of the largest E or H value, spectra 1e-4 of their maximum, as the
  single-simulation parity tests).
* The batch is the same arithmetic as fdtd_step: with the simulations' own
  tiles, graph replay, direct launches and each simulation run alone agree
  bit for bit in the fields, on the row-blocked path and on the general
  kernel (FDTD_PATH=general), including runs split by individual steps
  between batch runs, and on every stream layout (one stream, one or three
  side streams, the early side start).
* The autotuned tiles change only rounding (group-normalised 1e-5).
"""
import os

import numpy as np
import pytest

import fdtd_amd as G
import oracle as O
from conftest import small_params

pytestmark = pytest.mark.gpu

C0 = 2.99792458e8


def make_params(mod, shape, phi=0.0, steps=300):
    p = small_params(mod, n=shape[0], ncpml=6, sphere=0, steps=steps, phi=phi)
    p.NCELLX, p.NCELLY, p.NCELLZ = shape
    return p


# (interior shape, geometry, polarisation angle phi)
CASES = [
    ((24, 28, 26), ("sphere", 8e-9), 0.0),
    ((30, 24, 28), ("rod", 0, 22e-9, 5e-9), 90.0),    # rod along x, x-polarised
    ((26, 26, 32), ("rod", 1, 20e-9, 4e-9), 0.0),     # rod along y, y-polarised (E along -y at phi = 0)
    ((24, 24, 24), None, 90.0),                       # vacuum
]


def build(mod_sim, case, steps=300):
    shape, geo, phi = case
    p = make_params(O, shape, phi, steps)
    s = mod_sim(p)
    if geo is not None:
        T = [n + 2 * 6 + 1 for n in shape]
        c = [t // 2 for t in T]
        if geo[0] == "sphere":
            s.add_sphere(c, geo[1])
        else:
            s.add_rod(c, geo[1], geo[2], geo[3])
    return p, s


def sims_for(cases, env_general=False):
    old = os.environ.get("FDTD_PATH")
    if env_general:
        os.environ["FDTD_PATH"] = "general"
    try:
        return [build(G.Simulation, c)[1] for c in cases]
    finally:
        if env_general:
            if old is None:
                del os.environ["FDTD_PATH"]
            else:
                os.environ["FDTD_PATH"] = old


def oracle_for(case, sim, steps):
    o = O.Oracle(make_params(O, case[0], case[2]))
    o.set_materials(sim.get_materials(1), sim.get_materials(2), sim.get_materials(3))
    o.step(steps)
    return o


def field_errors(s, o):
    fo = [o.get_field(c) for c in range(6)]
    fs = [s.get_field(c) for c in range(6)]
    me = max(np.abs(f).max() for f in fo[:3])
    mh = max(np.abs(f).max() for f in fo[3:])
    return [float(np.abs(fs[c] - fo[c]).max() / (me if c < 3 else mh)) for c in range(6)]


def fields(s):
    return [s.get_field(c) for c in range(6)]


def test_rod_masks():
    """The rod is symmetric about its centre and its metal volume is that of a
    capped cylinder to within the staircase error."""
    case = CASES[1]
    _, s = build(G.Simulation, case)
    m = s.get_materials(0) > 1
    dx = 2e-9
    L, r = 22e-9, 5e-9
    vol = np.pi * r * r * (L - 2 * r) + 4.0 / 3.0 * np.pi * r ** 3
    cells = m.sum() * dx ** 3
    print("rod cells", m.sum(), "volume ratio", cells / vol, flush=True)
    assert abs(cells / vol - 1) < 0.25
    # centred: the mask is symmetric under reflection through the centre cell along y and z
    T = [n + 13 for n in case[0]]
    c = [t // 2 for t in T]
    sub = m[:, c[1] - 6:c[1] + 7, c[2] - 6:c[2] + 7]
    assert np.array_equal(sub, sub[:, ::-1, :]) and np.array_equal(sub, sub[:, :, ::-1])
    assert s.get_info().n_dispersive[0] > 0


@pytest.mark.parametrize("flags", [0, G.BATCH_GENERAL])
def test_batch_matches_oracle(flags):
    steps = 301      # not a multiple of the sampling period: direct steps after the graph replays
    sims = sims_for(CASES)
    b = G.Batch(sims, flags=flags)
    b.run(steps)
    for case, s in zip(CASES, sims):
        o = oracle_for(case, s, steps)
        errs = field_errors(s, o)
        fs, fo = s.flux(), o.flux()
        ext_err = float(np.max(np.abs(fs["ext"] - fo["ext"])) / max(np.max(np.abs(fo["ext"])), 1e-300))
        raw_err = float(np.max(np.abs(fs["raw"] - fo["raw"])) / np.max(np.abs(fo["raw"])))
        print(case[0], "tile", b.tile(CASES.index(case)), "fields", ["%.1e" % e for e in errs],
              "ext %.1e raw %.1e" % (ext_err, raw_err), flush=True)
        assert max(errs) < 2e-4, errs
        assert raw_err < 1e-4
        if case[1] is not None:
            assert ext_err < 1e-4


@pytest.mark.parametrize("general", [False, True])
def test_batch_graph_direct_and_alone_agree_bitwise(general):
    steps = 203
    a = sims_for(CASES, env_general=general)
    b = sims_for(CASES, env_general=general)
    c = sims_for(CASES, env_general=general)
    ba = G.Batch(a, flags=G.BATCH_NO_TUNE)
    bb = G.Batch(b, flags=G.BATCH_NO_TUNE | G.BATCH_NO_GRAPH)
    assert all(ba.tile(i)[0] == (0 if general else 1) for i in range(len(CASES)))
    ba.run(steps)
    bb.run(steps)
    for s in c:
        s.step(steps)
    for i in range(len(CASES)):
        fa, fb, fc = fields(a[i]), fields(b[i]), fields(c[i])
        for k in range(6):
            assert np.array_equal(fa[k], fb[k]), (i, k)
            assert np.array_equal(fa[k], fc[k]), (i, k)
        ra, rc = a[i].flux()["raw"], c[i].flux()["raw"]
        assert np.max(np.abs(ra - rc)) <= 1e-12 * np.max(np.abs(rc))
        assert a[i].get_info().t == c[i].get_info().t == steps
        assert a[i].get_info().sample_count == c[i].get_info().sample_count


@pytest.mark.parametrize("side", ["1", "3"])
def test_batch_early_side_start_agrees_bitwise(side, monkeypatch):
    """The early side start (FDTD_BATCH_EARLY=1: the next step's x and y
    slabs wait for the join only, and run beside the TFSF fix-ups,
    dispersive boxes and samples), on one side stream (the lines after the
    side pieces) and on three (FDTD_SIDE_STREAMS=3, the lines and corners on
    their own): graph replay, direct launches and each simulation alone agree
    bit for bit in the fields, spectra to rounding of the contraction.  Every
    simulation allows the start (its own step sets schedule bit 3 on the same
    cell conditions), so the batch takes it."""
    steps = 203
    a, b, c = sims_for(CASES), sims_for(CASES), sims_for(CASES)
    sched = [s.get_info().schedule for s in a]
    print("schedules", sched, flush=True)
    assert all(x & 8 for x in sched), sched
    monkeypatch.setenv("FDTD_SIDE_STREAMS", side)
    monkeypatch.setenv("FDTD_BATCH_EARLY", "1")
    ba = G.Batch(a, flags=G.BATCH_NO_TUNE)
    bb = G.Batch(b, flags=G.BATCH_NO_TUNE | G.BATCH_NO_GRAPH)
    ba.run(steps)
    bb.run(steps)
    for s in c:
        s.step(steps)
    for i in range(len(CASES)):
        fa, fb, fc = fields(a[i]), fields(b[i]), fields(c[i])
        for k in range(6):
            assert np.array_equal(fa[k], fb[k]), (i, k)
            assert np.array_equal(fa[k], fc[k]), (i, k)
        ra, rb, rc = a[i].flux()["raw"], b[i].flux()["raw"], c[i].flux()["raw"]
        assert np.max(np.abs(ra - rc)) <= 1e-12 * np.max(np.abs(rc))
        assert np.max(np.abs(rb - rc)) <= 1e-12 * np.max(np.abs(rc))


@pytest.mark.parametrize("env", [{"FDTD_CONCURRENT": "0"}, {"FDTD_SIDE_STREAMS": "3"}])
def test_batch_stream_layouts_agree_bitwise(env, monkeypatch):
    """Every batched launch on one stream (FDTD_CONCURRENT=0), and the side
    launches on three streams (FDTD_SIDE_STREAMS=3: x slabs, y slabs, lines
    and corner columns): graph replay, direct launches and each simulation
    alone agree bit for bit in the fields."""
    steps = 203
    a, b, c = sims_for(CASES), sims_for(CASES), sims_for(CASES)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    ba = G.Batch(a, flags=G.BATCH_NO_TUNE)
    bb = G.Batch(b, flags=G.BATCH_NO_TUNE | G.BATCH_NO_GRAPH)
    ba.run(steps)
    bb.run(steps)
    for s in c:
        s.step(steps)
    for i in range(len(CASES)):
        fa, fb, fc = fields(a[i]), fields(b[i]), fields(c[i])
        for k in range(6):
            assert np.array_equal(fa[k], fb[k]), (env, i, k)
            assert np.array_equal(fa[k], fc[k]), (env, i, k)
        ra, rc = a[i].flux()["raw"], c[i].flux()["raw"]
        assert np.max(np.abs(ra - rc)) <= 1e-12 * np.max(np.abs(rc))


@pytest.mark.parametrize("ctj", [8, 16])
def test_batch_corner_tiles_agree_bitwise(ctj, monkeypatch):
    """Corner columns of 8 and 16 tile rows (FDTD_CTJ; the batch groups its
    corner launches by tile rows): graph replay and each simulation alone
    agree bit for bit, as with the default 4."""
    monkeypatch.setenv("FDTD_CTJ", str(ctj))
    steps = 61
    a = sims_for(CASES)
    c = sims_for(CASES)
    ba = G.Batch(a, flags=G.BATCH_NO_TUNE)
    ba.run(steps)
    for s in c:
        s.step(steps)
    for i in range(len(CASES)):
        fa, fc = fields(a[i]), fields(c[i])
        for k in range(6):
            assert np.array_equal(fa[k], fc[k]), (ctj, i, k)


@pytest.mark.parametrize("general", [False, True])
def test_batch_runs_interleaved_with_single_steps(general):
    """Batch runs split by fdtd_step calls on every simulation (odd counts,
    so parities and the sampling phase move) equal one uninterrupted run."""
    a = sims_for(CASES, env_general=general)
    c = sims_for(CASES, env_general=general)
    ba = G.Batch(a, flags=G.BATCH_NO_TUNE)
    ba.run(97)
    for s in a:
        s.step(5)
    _ = [s.flux() for s in a]        # contracts a partial ring between runs
    ba.run(102)
    for s in c:
        s.step(204)
    for i in range(len(CASES)):
        fa, fc = fields(a[i]), fields(c[i])
        for k in range(6):
            assert np.array_equal(fa[k], fc[k]), (i, k)
        # the extra contraction of a partial ring regroups the FP32 partial sums of the DFT
        ra, rc = a[i].flux()["raw"], c[i].flux()["raw"]
        assert np.max(np.abs(ra - rc)) <= 1e-5 * np.max(np.abs(rc))


@pytest.mark.parametrize("flags", [0, G.BATCH_GENERAL])
def test_autotuned_tiles_change_only_rounding(flags):
    steps = 160
    a = sims_for(CASES)
    b = sims_for(CASES)
    ta = G.Batch(a, flags=flags)
    tb = G.Batch(b, flags=flags | G.BATCH_NO_TUNE)
    ta.run(steps)
    tb.run(steps)
    for i in range(len(CASES)):
        fa, fb = fields(a[i]), fields(b[i])
        me = max(np.abs(f).max() for f in fb[:3])
        mh = max(np.abs(f).max() for f in fb[3:])
        err = max(float(np.abs(fa[k] - fb[k]).max() / (me if k < 3 else mh)) for k in range(6))
        print("tile", ta.tile(i), "vs default: %.1e" % err, flush=True)
        assert err < 1e-5


def test_batch_rejects_mismatched_steps():
    sims = sims_for(CASES[:2])
    sims[0].step(3)
    with pytest.raises(G.FDTDError):
        G.Batch(sims)
