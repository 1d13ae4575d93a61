"""ctypes mirror of include/fdtd.h: the parameter and grid-info structs.

Field order must match the C header exactly; test_abi.py checks the struct
sizes against the compiled library.
"""
import ctypes as C

FDTD_OK = 0
FDTD_ERR_INVALID = -1
FDTD_ERR_ALLOC = -2
FDTD_ERR_DEVICE = -3
FDTD_ERR_NONFINITE = -4
FDTD_ERR_UNSUPPORTED = -5

EX, EY, EZ, HX, HY, HZ = range(6)


class Params(C.Structure):
    _fields_ = [
        ("Tend", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("NCELLX", C.c_int), ("NCELLY", C.c_int), ("NCELLZ", C.c_int),
        ("NcpmlX", C.c_int), ("NcpmlY", C.c_int), ("NcpmlZ", C.c_int),
        ("nano_sphere", C.c_int),
        ("nano_sphere_radius", C.c_double),
        ("material", C.c_int),
        ("first_medium", C.c_int),
        ("NUM_freq", C.c_int),
        ("WL_or_freq", C.c_int),
        ("BandWidth", C.c_double),
        ("NONLOCAL", C.c_int),
        ("MAX_AMP", C.c_double),
        ("TimeFactor", C.c_double),
        ("lam_0", C.c_double),
        ("StaticField", C.c_int),
        ("tfsf_inset", C.c_int),
        ("abs_offset", C.c_int),
        ("sca_offset", C.c_int),
        ("inc_theta", C.c_double), ("inc_phi", C.c_double), ("polar_psi", C.c_double),
        ("cpml_exp", C.c_int),
        ("kappa_max", C.c_double),
        ("alpha_max", C.c_double),
        ("sigma_factor", C.c_double),
        ("lam_min", C.c_double), ("lam_max", C.c_double),
        ("dft_stride", C.c_int),
        ("dft_block", C.c_int),
        ("source_offset", C.c_int),
        ("hydro_beta", C.c_double),
        ("xs_area", C.c_double),
        ("dft_exact", C.c_int),
        ("periodic_xy", C.c_int),
    ]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class GridInfo(C.Structure):
    _fields_ = [
        ("NX", C.c_int), ("NY", C.c_int), ("NZ", C.c_int),
        ("Ncpml", C.c_int * 3),
        ("dt", C.c_double),
        ("tfsf", C.c_int * 3),
        ("abs_start", C.c_int * 3), ("abs_end", C.c_int * 3),
        ("sca_start", C.c_int * 3), ("sca_end", C.c_int * 3),
        ("inc_length", C.c_int), ("m0", C.c_int),
        ("width", C.c_double), ("delay", C.c_double), ("f_0", C.c_double),
        ("num_freq", C.c_int),
        ("t", C.c_int),
        ("hydro", C.c_int),
        ("n_dispersive", C.c_int * 3),
        ("box_lo", C.c_int * 3), ("box_hi", C.c_int * 3),
        ("sample_count", C.c_int),
        ("fused_path", C.c_int),
        ("clean", C.c_int * 4),
        ("schedule", C.c_int),
        ("tile", C.c_int * 7),
    ]


def face_shape(info, face):
    """(na, nb) of monitor face 0..11 (box*6 + axis*2 + side)."""
    box, axis = face // 6, (face // 2) % 3
    s = info.abs_start if box else info.sca_start
    e = info.abs_end if box else info.sca_end
    outer = 1 if axis == 0 else 0
    inner = 1 if axis == 2 else 2
    return e[outer] - s[outer], e[inner] - s[inner]
