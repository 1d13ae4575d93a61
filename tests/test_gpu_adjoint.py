"""The GPU discrete adjoint (fdtd_set_design / fdtd_adjoint) against the FP64
oracle adjoint, whose own gradient is pinned by central differences in
test_oracle_adjoint.py.

FP32 on the GPU against FP64: J within 1e-3 relative, dJ/drho within 2e-3 of
the largest gradient entry.  Checkpoint spacing, and keeping the trajectory
instead of recomputing, must not change a single bit of the gradient (every
segment is recomputed by the same kernels), and the state left behind is the
plain forward state after nsteps.  The transposed step on the row-blocked
kernel (clean box, and its variants on the x and y slabs) agrees with the
general transposed kernel to rounding.
"""
import os

import numpy as np
import pytest

import fdtd_amd as G
import oracle as O
from conftest import small_params

C0 = 2.99792458e8
LAMS = np.array([633e-9, 785e-9, 850e-9])
STEPS = 260


def problem(mod, shape=(5, 4, 3), seed=3):
    p = small_params(mod, n=20, ncpml=5, sphere=0, steps=STEPS)
    p.inc_phi = 90.0
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.1, 0.9, size=shape)
    T = 20 + 2 * 5 + 1
    c = T // 2
    lo = (c - shape[0] // 2, c - shape[1] // 2, c - shape[2] // 2)
    r0 = (c, c + 1, lo[2] + shape[2] + 1)
    return p, rho, lo, r0


def oracle_adjoint(shape=(5, 4, 3)):
    p, rho, lo, r0 = problem(O, shape)
    o = O.Oracle(p)
    o.set_design(lo, rho.astype(np.float32).astype(np.float64))
    o.set_objective(r0, C0 / LAMS)
    return o.adjoint(STEPS)


@pytest.mark.gpu
def test_gpu_adjoint_matches_oracle():
    p, rho, lo, r0 = problem(G)
    sim = G.Simulation(p)
    sim.set_design(lo, rho)
    J, g = sim.adjoint(STEPS, r0, C0 / LAMS, checkpoint=50)
    Jo, go = oracle_adjoint()
    print("J", J, Jo, "max|g|", np.abs(go).max(), "err", np.abs(g - go).max(), flush=True)
    assert J == pytest.approx(Jo, rel=1e-3)
    assert np.abs(g - go).max() < 2e-3 * np.abs(go).max()


@pytest.mark.gpu
def test_gpu_adjoint_checkpoint_spacing_is_exact():
    p, rho, lo, r0 = problem(G)
    out = []
    for ck in (7, 64, STEPS, 0):
        sim = G.Simulation(p)
        sim.set_design(lo, rho)
        out.append(sim.adjoint(STEPS, r0, C0 / LAMS, checkpoint=ck))
        sim.close()
    for J, g in out[1:]:
        assert J == out[0][0]
        assert np.array_equal(g, out[0][1])


@pytest.mark.gpu
def test_gpu_adjoint_leaves_forward_state():
    p, rho, lo, r0 = problem(G)
    a = G.Simulation(p)
    a.set_design(lo, rho)
    a.adjoint(STEPS, r0, C0 / LAMS, checkpoint=40)
    b = G.Simulation(p)
    b.set_design(lo, rho)
    b.step(STEPS)
    for c in range(6):
        assert np.array_equal(a.get_field(c), b.get_field(c)), c
    fa, fb = a.flux(), b.flux()
    assert np.array_equal(fa["ext"], fb["ext"])
    # and the simulation carries on from there
    a.step(10)
    b.step(10)
    assert np.array_equal(a.get_field(G.EX), b.get_field(G.EX))


@pytest.mark.gpu
def test_gpu_adjoint_autograd_wrapper():
    torch = pytest.importorskip("torch")
    from fdtd_amd.design import DesignObjective
    p, rho, lo, r0 = problem(G)
    fn = DesignObjective()
    r = torch.tensor(rho, dtype=torch.float64, requires_grad=True)
    J = fn.apply(r, p, lo, r0, C0 / LAMS, STEPS, 32)
    (3.0 * J).backward()
    sim = G.Simulation(p)
    sim.set_design(lo, rho)
    J2, g2 = sim.adjoint(STEPS, r0, C0 / LAMS, checkpoint=32)
    assert float(J) == J2
    assert np.array_equal(r.grad.numpy(), 3.0 * g2)



def _adjoint_with_path(path, ck=40, shape=(5, 4, 3), slabs="rows", r0=None, extra=None):
    env = {"FDTD_ADJ_PATH": path, "FDTD_ADJ_SLABS": slabs, **(extra or {})}
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        p, rho, lo, r0d = problem(G, shape)
        sim = G.Simulation(p)
        sim.set_design(lo, rho)
        return sim.adjoint(STEPS, r0 or r0d, C0 / LAMS, checkpoint=ck)
    finally:
        for k, v in saved.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


@pytest.mark.gpu
@pytest.mark.parametrize("rows", ["2", "3"])
def test_gpu_adjoint_row_kernel_matches_general_kernel(rows):
    """Two or three rows per thread (FDTD_ROWS; the transposed clean box
    follows the forward step's, the slab variants keep two)."""
    old = os.environ.get("FDTD_ROWS")
    os.environ["FDTD_ROWS"] = rows
    try:
        info = G.Simulation(problem(G)[0]).get_info()
    finally:
        if old is None:
            del os.environ["FDTD_ROWS"]
        else:
            os.environ["FDTD_ROWS"] = old
    assert info.fused_path == 1 and info.tile[0] == int(rows)
    Jr, gr = _adjoint_with_path("rows", extra={"FDTD_ROWS": rows})
    Jg, gg = _adjoint_with_path("general", extra={"FDTD_ROWS": rows})
    err = np.abs(gr - gg).max() / np.abs(gg).max()
    print("rows %s: row vs general transposed step: max |dg| / max |g| = %.2e" % (rows, err), flush=True)
    assert Jr == Jg
    assert err < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("where", ["x slab", "y slab"])
def test_gpu_adjoint_slab_variants_match_general_kernel_and_oracle(where):
    """Objective cell inside the x-slab (or y-slab) CPML: the adjoint source
    starts in the slab, so the gradient on the block is carried there by the
    transposed slab kernels.  The row-blocked slab variants against the
    general transposed kernel on the slabs (same clean box) and on the whole
    grid, and against the FP64 oracle.  The objective there is the block's
    scattered field alone (about 3.5e-5 of the field at the default cell, so
    the FP32 TFSF leakage weighs more in it): 2e-2 against the oracle, 1e-5
    between the paths, whose forward passes are the same."""
    c = (20 + 2 * 5 + 1) // 2
    r0 = (3, c + 1, c) if where == "x slab" else (c, 3, c)
    Jr, gr = _adjoint_with_path("rows", r0=r0)
    Js, gs = _adjoint_with_path("rows", slabs="general", r0=r0)
    Jg, gg = _adjoint_with_path("general", r0=r0)
    po, rhoo, loo, _ = problem(O)
    o = O.Oracle(po)
    o.set_design(loo, rhoo.astype(np.float32).astype(np.float64))
    o.set_objective(r0, C0 / LAMS)
    Jo, go = o.adjoint(STEPS)
    scale = np.abs(gg).max()
    es, eg, eo = (np.abs(gr - x).max() / scale for x in (gs, gg, go))
    print("%s: slab rows vs slab general %.2e, vs general %.2e, vs oracle %.2e (max|g| %.3e)"
          % (where, es, eg, eo, scale), flush=True)
    assert scale > 0
    assert Jr == Js == Jg
    assert es < 1e-5 and eg < 1e-5
    assert Jr == pytest.approx(Jo, rel=2e-2)
    assert eo < 2e-2


@pytest.mark.gpu
@pytest.mark.parametrize("where", ["above the box", "in the box margin"])
def test_gpu_adjoint_schedules_give_the_same_bits(where):
    """The transposed dispersive update beside the clean box (schedules 1-3,
    3 with the correction amounts computed there and added after the clean
    box), the next step's slabs starting at the join (FDTD_ADJ_EARLY; the
    forward recompute's side pieces likewise, FDTD_EARLY_SIDE) and the next
    step's injection folded into the dispersive-cell correction
    (for an objective cell outside and inside the corrected cells) change no
    bit against the plain order: update, injection and correction on the
    main stream."""
    shape = (5, 4, 3)
    _, _, lo, r0 = problem(G, shape)
    if where == "in the box margin":
        r0 = (r0[0], r0[1], lo[2] + shape[2])
    J0, g0 = _adjoint_with_path("rows", r0=r0, extra={"FDTD_ADJ_SCHED": "0", "FDTD_ADJ_FUSE_INJECT": "0",
                                                      "FDTD_ADJ_EARLY": "0", "FDTD_EARLY_SIDE": "0"})
    for sched, fuse, early in (("3", "1", "1"), ("3", "1", "0"), ("3", "0", "1"), ("2", "1", "1"), ("1", "1", "1"),
                               ("1", "0", "1"), ("0", "1", "1")):
        J, g = _adjoint_with_path("rows", r0=r0, extra={"FDTD_ADJ_SCHED": sched, "FDTD_ADJ_FUSE_INJECT": fuse,
                                                        "FDTD_ADJ_EARLY": early, "FDTD_EARLY_SIDE": "2" if early == "1" else "0"})
        print(where, "schedule", sched, "fused injection", fuse, "early slabs", early, "max |dg|", np.abs(g - g0).max(),
              flush=True)
        assert J == J0
        assert np.array_equal(g, g0)
    assert np.abs(g0).max() > 0


@pytest.mark.gpu
def test_gpu_adjoint_larger_block_against_oracle():
    """A design block wider than a wave's z run and reaching close to the
    z CPML, trajectory kept (no recompute)."""
    shape = (6, 7, 7)
    p, rho, lo, r0 = problem(G, shape, seed=11)
    sim = G.Simulation(p)
    sim.set_design(lo, rho)
    J, g = sim.adjoint(STEPS, r0, C0 / LAMS, checkpoint=0)
    po, rhoo, loo, r0o = problem(O, shape, seed=11)
    o = O.Oracle(po)
    o.set_design(loo, rhoo.astype(np.float32).astype(np.float64))
    o.set_objective(r0o, C0 / LAMS)
    Jo, go = o.adjoint(STEPS)
    print("J", J, Jo, "err", np.abs(g - go).max() / np.abs(go).max(), flush=True)
    assert J == pytest.approx(Jo, rel=1e-3)
    assert np.abs(g - go).max() < 2e-3 * np.abs(go).max()


@pytest.mark.gpu
def test_gpu_adjoint_repeated_designs_keep_device_memory_flat():
    """An optimisation loop on one handle: a new design block (two sizes in
    turn, so the dispersive box is rebuilt) before every adjoint.  The box and
    adjoint buffers of the last design are freed before the next ones are
    allocated, so the device memory in use stays flat (before, each round
    added the twelve adjoint field arrays, about 17 MB on this 71^3 grid)."""
    import ctypes

    def in_use():
        # the HIP runtime the library itself mapped (another copy of the runtime,
        # such as one bundled with torch, would see a different device context)
        paths = [ln.split()[-1] for ln in open("/proc/self/maps") if "libamdhip64" in ln and "torch" not in ln]
        assert paths, "the library's HIP runtime is not mapped"
        hip = ctypes.CDLL(paths[0])
        free, total = ctypes.c_size_t(), ctypes.c_size_t()
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
        return total.value - free.value

    p = small_params(G, n=60, ncpml=5, sphere=0, steps=120)
    p.inc_phi = 90.0
    T = 60 + 2 * 5 + 1
    c = T // 2
    sim = G.Simulation(p)
    rng = np.random.default_rng(5)
    used = []
    try:
        for it in range(6):
            shape = (5, 4, 3) if it % 2 == 0 else (6, 6, 4)
            lo = (c - shape[0] // 2, c - shape[1] // 2, c - shape[2] // 2)
            r0 = (c, c + 1, lo[2] + shape[2] + 1)
            sim.set_design(lo, rng.uniform(0.1, 0.9, size=shape).astype(np.float32))
            J, g = sim.adjoint(40, r0, C0 / LAMS, checkpoint=20)
            assert np.isfinite(J) and np.all(np.isfinite(g))
            used.append(in_use())
        growth = max(used[2:]) - used[1]
        print("device memory in use after each design (MB):", [round(u / 2**20, 1) for u in used],
              "growth after the second design: %.1f MB" % (growth / 2**20), flush=True)
        assert growth < 8 * 2**20
    finally:
        sim.close()


@pytest.mark.gpu
@pytest.mark.parametrize("nz", [60, 130])
def test_gpu_adjoint_taller_grids_row_kernel_matches_general_kernel(nz):
    """z columns of two and three waves (NCELLZ 60 and 130 with a 5-cell
    CPML): the row-blocked transposed step (clean box and slab variants)
    against the general transposed kernel."""
    res = {}
    for path in ("rows", "general"):
        old = os.environ.get("FDTD_ADJ_PATH")
        os.environ["FDTD_ADJ_PATH"] = path
        try:
            p = small_params(G, n=20, ncpml=5, sphere=0, steps=STEPS)
            p.NCELLZ = nz
            p.inc_phi = 90.0
            T = [20 + 11, 20 + 11, nz + 11]
            shape = (5, 4, 3)
            # the block at the height it has in the 31^3 problem, which the pulse reaches within the window
            lo = (T[0] // 2 - 2, T[1] // 2 - 2, 14)
            r0 = (T[0] // 2, T[1] // 2 + 1, lo[2] + shape[2] + 1)
            sim = G.Simulation(p)
            rho = np.random.default_rng(13).uniform(0.1, 0.9, size=shape).astype(np.float32)
            sim.set_design(lo, rho)
            res[path] = sim.adjoint(STEPS, r0, C0 / LAMS, checkpoint=40)
            if path == "rows":
                assert sim.get_info().fused_path == 1
            sim.close()
        finally:
            if old is None:
                del os.environ["FDTD_ADJ_PATH"]
            else:
                os.environ["FDTD_ADJ_PATH"] = old
    (Jr, gr), (Jg, gg) = res["rows"], res["general"]
    err = np.abs(gr - gg).max() / np.abs(gg).max()
    print("NCELLZ %d: row vs general transposed step %.2e (max|g| %.3e)" % (nz, err, np.abs(gg).max()), flush=True)
    assert Jr == Jg and np.abs(gg).max() > 1e-8
    assert err < 1e-5


@pytest.mark.gpu
def test_gpu_adjoint_corner_variant_matches_general_kernel_and_oracle():
    """FDTD_ADJ_CORNER_ROWS=1, objective cell inside a corner column of the
    CPML (x and y both in the layer), so the gradient on the block is carried
    there by the transposed corner kernel: the row-blocked corner variant
    against the general transposed kernel on the corners and on the whole
    grid, and against the FP64 oracle (bars as for the slab variants)."""
    r0 = (3, 3, (20 + 2 * 5 + 1) // 2)
    Jr, gr = _adjoint_with_path("rows", r0=r0, extra={"FDTD_ADJ_CORNER_ROWS": "1"})
    Jc, gc = _adjoint_with_path("rows", r0=r0, extra={"FDTD_ADJ_CORNER_ROWS": "0"})
    Jg, gg = _adjoint_with_path("general", r0=r0)
    po, rhoo, loo, _ = problem(O)
    o = O.Oracle(po)
    o.set_design(loo, rhoo.astype(np.float32).astype(np.float64))
    o.set_objective(r0, C0 / LAMS)
    Jo, go = o.adjoint(STEPS)
    scale = np.abs(gg).max()
    ec, eg, eo = (np.abs(gr - x).max() / scale for x in (gc, gg, go))
    print("corner: rows vs general corners %.2e, vs general %.2e, vs oracle %.2e (max|g| %.3e)" % (ec, eg, eo, scale), flush=True)
    assert scale > 0
    assert Jr == Jc == Jg
    assert ec < 1e-5 and eg < 1e-5
    assert Jr == pytest.approx(Jo, rel=2e-2)
    assert eo < 2e-2

