"""Parameter sets of the benchmark workloads, shared by bench.py, the tests and
the scripts that write the FP64 golden fixtures.

Every builder takes the module whose `default_params()` fills an
`fdtd_params` (the HIP binding `fdtd_amd` or the oracle binding), so the same
numbers reach both solvers.  Common mapping choices: N is the interior edge
(the augmented grid is N + 2 Ncpml + 1, setup.cu:377-393); dt is half the
Courant limit (TimeFactor = sqrt(3), setup.cu:359); "x-polarised, travelling
along z" is phi = 90 deg (tfsf2.cu:11-13 with the angles exposed); wavelength
lists are linear in wavelength (main.cu:2451).
"""
import math

C0 = 2.99792458e8


def plain_method(mod):
    """64^3 interior at dx = 2 nm, 10-cell CPML (m = 3, kappa_max = 5,
    alpha_max = 0.05), TFSF 8 cells inside the CPML, scattering box 4 cells
    outside it, one 20 nm gold sphere at the centre, 1000 steps, 16
    wavelengths 400-900 nm, 600 nm pulse with a 300 nm FWHM bandwidth
    (BandWidth 6.79e14 Hz)."""
    p = mod.default_params()
    p.NCELLX = p.NCELLY = p.NCELLZ = 64
    p.NcpmlX = p.NcpmlY = p.NcpmlZ = 10
    p.dx = p.dy = p.dz = 2e-9
    p.cpml_exp, p.kappa_max, p.alpha_max = 3, 5.0, 0.05
    p.tfsf_inset = 8
    p.sca_offset = 4
    p.abs_offset = 5
    p.inc_phi = 90.0
    p.TimeFactor = 3 ** 0.5
    p.lam_0 = 600e-9
    p.BandWidth = 6.79e14
    p.nano_sphere = 1
    p.nano_sphere_radius = 20e-9
    p.material = 3
    p.Tend = 1000
    p.NUM_freq = 16
    p.lam_min, p.lam_max = 400e-9, 900e-9
    p.dft_block = 64
    return p


# (int)(r/dx) truncates 30e-9/1e-9 = 29.999999999999996 to 29, so the radius
# handed to the staircase rule is 30.5 nm: 30-cell spheres, 2-cell gap.
DIMER_RADIUS = 30.5e-9
DIMER_STEPS = 20000


def dimer(mod, steps=DIMER_STEPS):
    """Precision-study dimer: 256^3 interior at dx = 1 nm, 12-cell CPML
    (281^3 cells), two 30 nm gold spheres with a 2 nm gap along x (placed by
    dimer_centres / DIMER_RADIUS, the grid has no sphere of its own),
    x-polarised plane wave along z (600 nm pulse, BandWidth 6.79e14 Hz), 64
    wavelengths 400-1000 nm buffered K = 64 samples per contraction, cross
    sections normalised by both spheres' geometric area."""
    p = mod.default_params()
    p.NCELLX = p.NCELLY = p.NCELLZ = 256
    p.NcpmlX = p.NcpmlY = p.NcpmlZ = 12
    p.dx = p.dy = p.dz = 1e-9
    p.nano_sphere = 0
    p.nano_sphere_radius = 30e-9
    p.material = 3
    p.Tend = steps
    p.NUM_freq = 64
    p.lam_min, p.lam_max = 400e-9, 1000e-9
    p.dft_block = 64
    p.inc_phi = 90.0
    p.TimeFactor = 3 ** 0.5
    p.tfsf_inset = 10
    p.xs_area = 2 * math.pi * (30e-9) ** 2
    return p


def dimer_centres(T):
    """Sphere centres (cells) of the dimer on a T^3 grid: 31 cells either side
    of the centre along x."""
    c = T // 2
    return [(c - 31, c, c), (c + 31, c, c)]


def dimer_gap(T):
    """Cell at the centre of the dimer's gap (the gap-field point)."""
    c = T // 2
    return (c, c, c)


HYDRO_STEPS = 40000


def hydro(mod, n=384, dx=0.25e-9, steps=HYDRO_STEPS):
    """Hydrodynamic nonlocal gold: n^3 interior, 12-cell CPML, one 10 nm
    sphere, beta = 1.27e6 m/s with the Drude + 2 CP poles, 48 wavelengths
    400-900 nm, the 600 nm pulse.  n = 384 at dx = 0.25 nm is the full case;
    n = 96 at dx = 1 nm the down-scaled copy (10-cell sphere) that the FP64
    oracle runs."""
    p = mod.default_params()
    p.NCELLX = p.NCELLY = p.NCELLZ = n
    p.NcpmlX = p.NcpmlY = p.NcpmlZ = 12
    p.dx = p.dy = p.dz = dx
    p.TimeFactor = 3 ** 0.5
    p.inc_phi = 90.0
    p.nano_sphere = 1
    p.nano_sphere_radius = 10e-9
    p.material = 3
    p.NONLOCAL = 3
    p.hydro_beta = 1.27e6
    p.tfsf_inset = 10
    p.Tend = steps
    p.NUM_freq = 48
    p.lam_min, p.lam_max = 400e-9, 900e-9
    p.dft_block = 64
    return p


DESIGN_BLOCK = (64, 64, 32)
DESIGN_STEPS = 10000
DESIGN_CHECKPOINT = 500
DESIGN_LAMBDAS = (633e-9, 785e-9, 850e-9)


def design(mod, n=256, steps=DESIGN_STEPS):
    """Inverse design: n^3 interior at dx = 2 nm, 12-cell CPML, 600 nm
    x-polarised pulse along z, 3 wavelengths 633-850 nm; the density block
    and objective point come from design_geometry."""
    p = mod.default_params()
    p.NCELLX = p.NCELLY = p.NCELLZ = n
    p.NcpmlX = p.NcpmlY = p.NcpmlZ = 12
    p.dx = p.dy = p.dz = 2e-9
    p.nano_sphere = 0
    p.material = 3
    p.Tend = steps
    p.NUM_freq = 3
    p.lam_min, p.lam_max = 633e-9, 850e-9
    p.inc_phi = 90.0
    p.TimeFactor = 3 ** 0.5
    p.tfsf_inset = 10
    p.lam_0 = 600e-9
    p.BandWidth = 6.79e14
    return p


def design_geometry(T, block=DESIGN_BLOCK):
    """(lo, r0) on a T^3 grid: the block centred in the grid, and the
    objective cell 5 nm (2.5 cells, rounded to 2 cells past the last block
    plane, i.e. the cell centre 2.5 cells above the top face) above the
    block's top face on its axis."""
    lo = tuple(T // 2 - k // 2 for k in block)
    r0 = (T // 2, T // 2, lo[2] + block[2] + 2)
    return lo, r0


def design_freqs():
    return [C0 / lam for lam in DESIGN_LAMBDAS]
