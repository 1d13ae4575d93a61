"""Throughput benchmark of the FDTD hot path on one MI355X.

Workload (the 256^3 precision-study configuration): 256^3 interior cells at
dx = 1 nm with a 12-cell CPML on every side (281^3 cells in total), a gold
dimer of two 30 nm spheres with a 2 nm gap along x (Drude + 2 critical-point
model, staircased), an x-polarised plane wave travelling along z through a
TFSF box, and DFT flux monitors at 64 wavelengths from 400 to 1000 nm buffered
K = 64 samples per contraction.  One benchmark step is one full FDTD time
step: 1D line, fused H/CPML/TFSF/E update, dispersive ADE, the DFT sample
every 4th step, and 1/256 of a K = 64-sample contraction (the contraction runs
once per 64 samples = 256 steps).

Prints one JSON line.  value = Mcell-updates/s = 281^3 / ms_per_step, where
ms_per_step is the wall time of the timed steps (inputs resident in HBM) with
the contractions that fell inside the window replaced by the amortised share
of one contraction timed in the same run (HIP events): a window of any length
then costs every step its full work.  roofline: the dominant kernel's
algorithmic bytes per launch / its average launch time from HIP events.
cpu_baseline: the FP64 OpenMP oracle on a bounded sample of the same grid.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("fdtd-cuda_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

HBM_PEAK_GBS = 8000.0


from fdtd_amd import workloads as W  # noqa: E402

workload_params = W.dimer
dimer_centres = W.dimer_centres
DIMER_RADIUS = W.DIMER_RADIUS


def algorithmic_bytes(info, p):
    """FP32 bytes one fused step must move over the whole grid.

    Each of the 6 field components is read once and written once (48 B per
    cell).  Cells inside a CPML slab add 4 psi values (2 H + 2 E components
    with a derivative along that axis) read and written per axis: 32 B.
    """
    N = np.array([info.NX, info.NY, info.NZ], dtype=np.int64)
    nc = np.array(info.Ncpml[:], dtype=np.int64)
    cells = int(N.prod())
    total = 48 * cells
    for a in range(3):
        others = int(np.prod([N[b] for b in range(3) if b != a]))
        total += 32 * 2 * int(nc[a]) * others
    return total


def clean_bytes(info):
    """Algorithmic bytes of one clean-box launch of the row-blocked kernel:
    48 B per cell of x [c0, c1) by y [c2, c3) by all z, plus 16 B per z-CPML
    psi pair (H: hx, hy at k < Nc and k >= F; E: ex, ey at 0 < k <= Nc and
    k >= F; each read and written)."""
    c = list(info.clean)
    if c[0] < 0:
        return None
    NZ, nc = info.NZ, info.Ncpml[2]
    F = NZ - 1 - nc
    area = (c[1] - c[0]) * (c[3] - c[2])
    psi_h = nc + max(0, NZ - 1 - F)
    psi_e = nc + max(0, NZ - 1 - F)
    return area * (48 * NZ + 16 * (psi_h + psi_e))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2048)
    ap.add_argument("--warmup", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sweep-steps", type=int, default=128,
                    help="steps of the 32-rod sweep window, batched and one by one (0: skip)")
    ap.add_argument("--adjoint-steps", type=int, default=10000,
                    help="steps of the adjoint/forward timing on the inverse-design grid (0: skip)")
    ap.add_argument("--hydro-steps", type=int, default=200,
                    help="steps of the hydrodynamic 384^3 timing window (0: skip)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 1 GiB -> 1 GiB copy kernel 4 times (reference for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs each simulation on a single MI355X "
              "(multi-GPU slab decomposition is not implemented); rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import fdtd_amd

    p = workload_params(fdtd_amd)
    sim = fdtd_amd.Simulation(p)
    info = sim.info
    T = info.NX
    for c in dimer_centres(T):
        sim.add_sphere(c, DIMER_RADIUS)
    info = sim.get_info()
    cells = info.NX * info.NY * info.NZ

    sim.step(args.warmup)
    sim.sync()
    rows = info.fused_path == 1
    cbytes = clean_bytes(info)
    K, stride = p.dft_block, p.dft_stride
    # timed run: events around the dominant launch of every step (the clean
    # box of the row-blocked kernel; the whole update when there is none)
    sim.kernel_timing(2 if cbytes else 1)
    n0 = sim.get_info().sample_count
    t0 = time.perf_counter()
    sim.step(args.steps)
    sim.sync()
    wall = time.perf_counter() - t0
    n1 = sim.get_info().sample_count
    # contractions that ran inside the window: one each time the ring filled
    in_window = n1 // K - n0 // K
    k_ms, k_n = sim.kernel_timing(1)
    # a shorter run timing all field-update kernels of a step together
    nu = max(16, args.steps // 4)
    sim.step(nu)
    u_ms, u_n = sim.kernel_timing(3)
    # one K-sample contraction (twiddles + contraction kernel), HIP events
    sim.step(K * stride)
    c_ms, c_n = sim.kernel_timing(0)
    contraction_ms = c_ms / max(c_n, 1)
    sim.sync()

    ms_step = (wall * 1e3 - in_window * contraction_ms) / args.steps + contraction_ms / (K * stride)
    value = cells / (ms_step * 1e-3) / 1e6
    abytes = algorithmic_bytes(info, p)
    kbytes = cbytes if cbytes else abytes
    k_avg_s = (k_ms / 1e3) / max(k_n, 1)
    achieved = kbytes / k_avg_s / 1e9
    u_avg_s = (u_ms / 1e3) / max(u_n, 1)

    sim.close()
    # the same kernel with the side pieces on the main stream after it (one
    # stream), so its duration is not shared with them: reported beside the
    # roofline figure, which is the production two-stream schedule
    serial = serial_kernel_ms(p, max(64, args.steps // 8)) if cbytes else None
    if args.calibrate_copy:
        copy_calibration()
    # the 32-simulation sweep last: after its allocations and frees the device
    # memory is left in a state that slows later large simulations of the same
    # process by 8-10 % (measured on the adjoint), which the dimer run does not
    adjoint = adjoint_bench(args.adjoint_steps) if args.adjoint_steps > 0 else None
    hydro = hydro_bench(args.hydro_steps) if args.hydro_steps > 0 else None
    sweep = sweep_bench(args.sweep_steps) if args.sweep_steps > 0 else None

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(p, info, args.cpu_steps)

    out = {
        "metric": "Mcell-updates/s",
        "value": round(value, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "wall_ms_per_step": round(wall * 1e3 / args.steps, 4),
        "contraction": {"kernels": "k_twiddle_frag + k_dft_mfma (split-BF16 MFMA, K=%d)" % K,
                        "ms": round(contraction_ms, 4), "every_steps": K * stride, "in_window": in_window,
                        "amortised_ms_per_step": round(contraction_ms / (K * stride), 5)},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "gold dimer 2x30 nm, 2 nm gap, 256^3 interior + 12-cell CPML (281^3), dx=1 nm, "
                        "TFSF x-pol plane wave, Drude+2CP, 64-wavelength DFT flux monitors (K=64)",
            "grid": [info.NX, info.NY, info.NZ],
            "dispersive_cells": list(info.n_dispersive[:]),
            # the tiles fdtd_step chose (fdtd_grid_info.tile)
            "tile": dict(zip(["rows", "cchunk", "schunk", "row_tiles", "side_row_tiles", "corner_tj", "side_streams"],
                             list(info.tile)), schedule=info.schedule),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_step_rows<clean>" if cbytes else ("k_step_rows" if rows else "k_fused_step"),
            "clean_box": list(info.clean),
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_per_launch(),
            "bytes_per_launch": kbytes,
            "kernel_avg_ms": round(k_avg_s * 1e3, 5),
            "launches": k_n,
            "one_stream_kernel_avg_ms": round(serial, 5) if serial else None,
            "one_stream_frac": round(kbytes / (serial / 1e3) / 1e9 / HBM_PEAK_GBS, 4) if serial else None,
        },
        "field_update": {
            # the window runs from the fork to the join of the two streams: it
            # also holds the 1D line and, with schedule bits 1 and 2, the
            # dispersive box and the DFT sample (every fourth step), whose bytes
            # are not counted (the figure is a lower bound)
            "kernels": ("k_line + k_step_rows + k_fused_step (corners) + "
                        + ("k_post (TFSF fix-up" + (" + dispersive box" if info.schedule & 2 else "")
                           + (" + DFT sample" if info.schedule & 4 else "") + ")"
                           if info.schedule & 6 else "k_tfsf_fix")) if rows else "k_fused_step",
            "avg_ms": round(u_avg_s * 1e3, 5),
            "bytes_per_step": abytes,
            "achieved_GBs": round(abytes / u_avg_s / 1e9, 1),
            "frac": round(abytes / u_avg_s / 1e9 / HBM_PEAK_GBS, 4),
        },
        "sweep": sweep,
        "adjoint": adjoint,
        "hydro": hydro,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def serial_kernel_ms(p, steps):
    """Average HIP-event duration of the clean-box launch with every kernel of
    the step on one stream (FDTD_CONCURRENT=0), on the same workload."""
    import fdtd_amd
    old = os.environ.get("FDTD_CONCURRENT")
    os.environ["FDTD_CONCURRENT"] = "0"
    try:
        sim = fdtd_amd.Simulation(p)
    finally:
        if old is None:
            del os.environ["FDTD_CONCURRENT"]
        else:
            os.environ["FDTD_CONCURRENT"] = old
    for c in dimer_centres(sim.info.NX):
        sim.add_sphere(c, DIMER_RADIUS)
    sim.step(16)
    sim.kernel_timing(2)
    sim.step(steps)
    ms, n = sim.kernel_timing(1)
    sim.close()
    return ms / max(n, 1)


def traffic_per_launch():
    """HBM bytes per launch of the dominant kernel from the committed PMC
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor), or
    None when no such measurement is on file."""
    path = os.path.join(ROOT, "profiles", "k_step_rows_dimer281_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["traffic_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def copy_calibration(nbytes=1 << 30, reps=4):
    """A plain device-to-device copy of nbytes (read nbytes, write nbytes; the
    HIP runtime runs it as its copyBuffer kernel) so the FETCH_SIZE /
    WRITE_SIZE counters can be scaled against a known size."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    for _ in range(reps):
        if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3):   # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def sweep_bench(steps, seed=20240601):
    """Batched sweep of 32 seeded nanorod simulations of uneven sizes: a
    window of `steps` steps run one simulation at a time (each on its own
    fastest path, fdtd_step) and then batched (fdtd_batch_run, graph replay),
    as aggregate Mcell-updates/s over all cells of all simulations."""
    import fdtd_amd
    from fdtd_amd import sweep
    steps = max(4, steps // 4 * 4)
    cfgs, redraws = sweep.draw(seed)
    t0 = time.perf_counter()
    sims = [sweep.make_sim(fdtd_amd.Simulation, fdtd_amd, c) for c in cfgs]
    t_create = time.perf_counter() - t0
    cells = sweep.total_cells(cfgs)
    for s in sims:
        s.step(8)
        s.sync()
    t_alone = 0.0
    for s in sims:
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        t_alone += time.perf_counter() - t0
    t0 = time.perf_counter()
    b = fdtd_amd.Batch(sims)
    t_tune = time.perf_counter() - t0
    b.run(8)                      # warm-up, graph capture
    t0 = time.perf_counter()
    b.run(steps)
    t_batch = time.perf_counter() - t0
    tiles = sorted({b.tile(i) for i in range(len(sims))})
    b.close()
    for s in sims:
        s.close()
    return {
        "workload": "32 seeded gold nanorods (splitmix64 seed %d), edges 96..192 step 8, dx=1.5 nm, 10-cell CPML, "
                    "Drude+2CP, 32 wavelengths 500-1200 nm" % seed,
        "sims": len(cfgs),
        "redraws": redraws,
        "cells": cells,
        "steps": steps,
        "batched_mcells_per_s": round(cells * steps / t_batch / 1e6, 1),
        "one_by_one_mcells_per_s": round(cells * steps / t_alone / 1e6, 1),
        "speedup": round(t_alone / t_batch, 3),
        "ms_per_step_batched": round(t_batch * 1e3 / steps, 3),
        "tiles_path_a_c": [list(t) for t in tiles],   # path 1: rows per thread, x chunk; 0: block rows, x chunk
        "autotune_s": round(t_tune, 2),
        "create_s": round(t_create, 2),
    }


def adjoint_bench(steps, ckpt=None, seed=7):
    """Adjoint/forward wall-time ratio on the inverse-design grid: 256^3
    interior at dx = 2 nm (281^3 cells), 12-cell CPML, a 64x64x32 density
    block at the centre with rho ~ U[0, 1) (seeded), objective |E|^2 at 633,
    785 and 850 nm at the cell 5 nm above the block's top face, x-polarised
    plane wave along z (600 nm centre, BandWidth 6.79e14 Hz; the pulse
    reaches the block after about 4000 steps, so J and dJ/drho are zero for
    shorter windows).  Forward = fdtd_step over `steps`; adjoint =
    fdtd_adjoint over the same steps (checkpoints every `ckpt` steps with
    recompute, then the stored trajectory)."""
    import fdtd_amd
    ckpt = ckpt or (W.DESIGN_CHECKPOINT if steps >= 2000 else max(1, steps // 4))
    p = W.design(fdtd_amd)
    T = 256 + 25
    n = W.DESIGN_BLOCK
    lo, r0 = W.design_geometry(T)
    rho = np.random.default_rng(seed).uniform(0.0, 1.0, size=n)
    freqs = np.array(W.design_freqs())
    f = fdtd_amd.Simulation(p)
    f.set_design(lo, rho)
    f.step(4)
    f.sync()
    t0 = time.perf_counter()
    f.step(steps)
    f.sync()
    t_fwd = time.perf_counter() - t0
    f.close()
    res = {}
    for mode, ck in (("recompute", ckpt), ("stored", 0)):
        a = fdtd_amd.Simulation(p)
        a.set_design(lo, rho)
        t0 = time.perf_counter()
        J, g = a.adjoint(steps, r0, freqs, checkpoint=ck)
        res[mode] = (time.perf_counter() - t0, J, float(np.linalg.norm(g)))
        a.close()
    t_rec, t_sto = res["recompute"][0], res["stored"][0]
    return {
        "workload": "inverse design: 256^3 interior dx=2 nm, 64x64x32 density block, |E|^2 at 633/785/850 nm",
        "steps": steps,
        "checkpoint": ckpt,
        "forward_ms_per_step": round(t_fwd * 1e3 / steps, 4),
        "stored_trajectory_ms_per_step": round(t_sto * 1e3 / steps, 4),
        "ratio_stored_trajectory": round(t_sto / t_fwd, 3),
        "checkpoint_recompute_ms_per_step": round(t_rec * 1e3 / steps, 4),
        "ratio_checkpoint_recompute": round(t_rec / t_fwd, 3),
        "J": res["recompute"][1],
        "grad_norm": res["recompute"][2],
        "stored_equals_recompute": res["stored"][1] == res["recompute"][1] and res["stored"][2] == res["recompute"][2],
    }


def hydro_bench(steps):
    """Hydrodynamic nonlocal gold: 384^3 interior at dx = 0.25 nm (409^3 cells),
    12-cell CPML, one 10 nm sphere (40 cells), beta = 1.27e6 m/s with the
    Drude + 2 CP poles, 48 wavelengths 400-900 nm; Mcell-updates/s over a
    window of `steps` full steps (velocity, charge and E-current kernels on the
    metal box after the field update)."""
    import fdtd_amd
    p = W.hydro(fdtd_amd)
    s = fdtd_amd.Simulation(p)
    info = s.get_info()
    cells = info.NX * info.NY * info.NZ
    s.step(16)
    s.sync()
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    wall = time.perf_counter() - t0
    s.close()
    return {
        "workload": "hydrodynamic gold sphere r=10 nm, 384^3 interior + 12-cell CPML (409^3), dx=0.25 nm, "
                    "beta=1.27e6 m/s, 48 wavelengths",
        "cells": cells,
        "metal_cells": int(info.n_dispersive[0]),
        "steps": steps,
        "ms_per_step": round(wall * 1e3 / steps, 4),
        "mcells_per_s": round(cells * steps / wall / 1e6, 1),
    }


def cpu_baseline(p, info, nsteps):
    """FP64 OpenMP oracle on the same 281^3 dimer grid: 2 warm-up steps, then nsteps timed."""
    import oracle as O
    o = O.Oracle(p)
    T = info.NX
    mx, my, mz = (o.get_materials(c) for c in (1, 2, 3))
    # the dimer masks exactly as the GPU solver built them
    import fdtd_amd
    sim = fdtd_amd.Simulation(p)
    for c in dimer_centres(T):
        sim.add_sphere(c, DIMER_RADIUS)
    o.set_materials(sim.get_materials(1), sim.get_materials(2), sim.get_materials(3))
    sim.close()
    del mx, my, mz
    o.step(2)
    t0 = time.perf_counter()
    o.step(nsteps)
    wall = time.perf_counter() - t0
    cells = info.NX * info.NY * info.NZ
    return {
        "value": round(cells * nsteps / wall / 1e6, 2),
        "unit": "Mcell-updates/s",
        "cores": o.threads(),
        "kind": "port",
        "sample": f"{nsteps} FP64 steps of the same 281^3 dimer grid after 2 warm-up steps ({wall:.1f} s)",
    }


if __name__ == "__main__":
    main()
