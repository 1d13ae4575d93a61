"""Seeded nanorod parameter sweep: many independent simulations of uneven
grid sizes, the workload of the batched path (fdtd_batch_*).

Each draw: interior edges NX, NY, NZ uniform in {96, 104, ..., 192}; one gold
nanorod with hemispherical caps, length uniform in [40, 120] nm, diameter
uniform in [10, 30] nm, axis uniform in {x, y, z}; dx = 1.5 nm, 10-cell CPML,
Drude + 2 critical-point gold, 32 DFT wavelengths from 500 to 1200 nm.  The
plane wave is polarised along the rod axis: it travels along z for x and y
rods and along x for z rods.  A draw whose rod does not fit inside the
absorption box with a 2-cell margin on every side is drawn again (the number
of redraws is reported).

Random numbers come from splitmix64 (Steele, Lea and Flood, "Fast splittable
pseudorandom number generators", OOPSLA 2014) mapped to [0, 1) as
(u >> 11) * 2^-53, so the draws are the same on every platform.
"""
import numpy as np

MASK64 = (1 << 64) - 1
EDGES = list(range(96, 193, 8))
DX = 1.5e-9
NCPML = 10


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK64

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next_u64() >> 11) * 2.0 ** -53


def _box(params_like, n_interior):
    """Absorption box [lo, hi) along one axis of the augmented grid."""
    N = n_interior + 2 * NCPML + 1
    tf0 = params_like["tfsf_inset"] + NCPML
    return tf0 + params_like["abs_offset"], N - tf0 - params_like["abs_offset"], N


def rod_fits(edges, axis, length, diameter, geom):
    for a in range(3):
        lo, hi, N = _box(geom, edges[a])
        c = N // 2
        half = (length if a == axis else diameter) / DX / 2.0
        # dispersive cells strictly inside the box, plus a 2-cell margin
        if c - half < lo + 1 + 2 or c + half > hi - 2 - 2:
            return False
    return True


def draw(seed=20240601, count=32, geom=None):
    """Returns (configs, redraws).  Each config: edges (3 interior sizes),
    axis, length and diameter (metres)."""
    if geom is None:
        geom = {"tfsf_inset": 10, "abs_offset": 5}
    rng = SplitMix64(seed)
    out, redraws = [], 0
    while len(out) < count:
        edges = [EDGES[min(int(rng.uniform() * len(EDGES)), len(EDGES) - 1)] for _ in range(3)]
        length = (40.0 + 80.0 * rng.uniform()) * 1e-9
        diameter = (10.0 + 20.0 * rng.uniform()) * 1e-9
        axis = min(int(rng.uniform() * 3), 2)
        if not rod_fits(edges, axis, length, diameter, geom):
            redraws += 1
            continue
        out.append({"edges": edges, "axis": axis, "length": length, "diameter": diameter})
    return out, redraws


def params_for(mod, cfg, steps=8000):
    """Solver parameters of one draw (mod: fdtd_amd or the oracle binding)."""
    p = mod.default_params()
    p.NCELLX, p.NCELLY, p.NCELLZ = cfg["edges"]
    p.NcpmlX = p.NcpmlY = p.NcpmlZ = NCPML
    p.dx = p.dy = p.dz = DX
    p.nano_sphere = 0
    p.material = 3
    p.Tend = steps
    p.NUM_freq = 32
    p.lam_min, p.lam_max = 500e-9, 1200e-9
    p.lam_0 = 800e-9
    p.dft_block = 64
    p.TimeFactor = 3 ** 0.5
    p.tfsf_inset = 10
    p.abs_offset = 5
    if cfg["axis"] == 2:
        p.inc_theta, p.inc_phi, p.polar_psi = 90.0, 0.0, 90.0    # along +x, E along +z
    else:
        p.inc_theta, p.polar_psi = 0.0, 0.0                       # along +z
        p.inc_phi = 90.0 if cfg["axis"] == 0 else 0.0             # E along +x, or along -y
    # geometric cross section seen along the propagation direction (side view of the capped rod)
    r = cfg["diameter"] / 2
    p.xs_area = 2 * r * (cfg["length"] - 2 * r) + np.pi * r * r
    return p


def make_sim(sim_cls, mod, cfg, steps=8000):
    s = sim_cls(params_for(mod, cfg, steps))
    c = [n // 2 for n in s.shape]
    s.add_rod(c, cfg["axis"], cfg["length"], cfg["diameter"] / 2)
    return s


def total_cells(cfgs):
    return int(sum(np.prod([e + 2 * NCPML + 1 for e in c["edges"]]) for c in cfgs))
