"""ctypes binding of the FP64 oracle (liboracle.so) — test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, always as the checker of the HIP solver, never as a compute path.
"""
import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.join(os.path.dirname(_HERE), "fdtd-cuda_amd")
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)

from fdtd_amd.abi import GridInfo, Params, face_shape  # noqa: E402

LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle not built: {LIB_PATH} (run `make -C oracle`)")
    lib = C.CDLL(LIB_PATH)
    vp, ip, dp, u8p = C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_uint8)
    sig = {
        "orc_last_error": (C.c_char_p, []),
        "orc_default_params": (None, [C.POINTER(Params)]),
        "orc_create": (ip, [C.POINTER(Params), C.POINTER(vp)]),
        "orc_destroy": (None, [vp]),
        "orc_info": (ip, [vp, C.POINTER(GridInfo)]),
        "orc_set_materials": (ip, [vp, u8p, u8p, u8p]),
        "orc_get_materials": (ip, [vp, ip, u8p]),
        "orc_add_sphere": (ip, [vp, dp, C.c_double]),
        "orc_get_state": (ip, [vp, ip, ip, dp]),
        "orc_add_layer": (ip, [vp, ip, ip]),
        "orc_reflectance": (ip, [vp, dp, dp, dp, dp]),
        "orc_update_h": (ip, [vp]), "orc_tfsf": (ip, [vp]), "orc_update_e": (ip, [vp]),
        "orc_end_step": (ip, [vp]), "orc_step": (ip, [vp, ip]),
        "orc_get_field": (ip, [vp, ip, dp]), "orc_set_field": (ip, [vp, ip, dp]),
        "orc_get_line": (ip, [vp, dp, dp]),
        "orc_get_spectrum_setup": (ip, [vp, dp, dp]),
        "orc_get_dft": (ip, [vp, ip, ip, dp]),
        "orc_flux": (ip, [vp, dp, dp, dp, dp]),
        "orc_coeffs": (ip, [vp, dp]),
        "orc_cpml": (ip, [vp, ip, ip, dp]),
        "orc_threads": (ip, []),
        "orc_set_design": (ip, [vp, C.POINTER(C.c_int * 3), C.POINTER(C.c_int * 3), dp]),
        "orc_set_objective": (ip, [vp, C.POINTER(C.c_int * 3), ip, dp]),
        "orc_objective": (C.c_double, [vp]),
        "orc_objective_terms": (C.c_int, [vp, C.POINTER(C.c_double)]),
        "orc_adjoint": (ip, [vp, ip, dp, dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def default_params():
    p = Params()
    load().orc_default_params(C.byref(p))
    return p


def _ptr(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


# names of orc_coeffs entries, in order
COEFF_NAMES = (["dt", "C_E", "C_E_1", "C_E_2", "d1", "d2", "d3", "d4", "d5", "dNL"]
               + [f"{n}_{i}" for i in range(2) for n in ("c1", "c2", "c3", "c4", "c5")]
               + ["eps_inf", "N_EQ", "Che", "Ceh"])
CPML_NAMES = ["be_N", "ce_N", "be_F", "ce_F", "bh_N", "ch_N", "bh_F", "ch_F", "ked", "khd"]


class Oracle:
    def __init__(self, params):
        self.lib = load()
        self.h = C.c_void_p()
        rc = self.lib.orc_create(C.byref(params), C.byref(self.h))
        if rc != 0:
            raise RuntimeError(f"oracle create failed: {self.lib.orc_last_error().decode()}")
        self.params = params
        self.info = self.get_info()
        self.shape = (self.info.NX, self.info.NY, self.info.NZ)

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_info(self):
        g = GridInfo()
        self.lib.orc_info(self.h, C.byref(g))
        return g

    def set_materials(self, mx, my, mz):
        arrs = [np.ascontiguousarray(m, dtype=np.uint8).reshape(-1) for m in (mx, my, mz)]
        self.lib.orc_set_materials(self.h, *[_ptr(a, C.c_uint8) for a in arrs])

    def add_sphere(self, centre, radius):
        c = np.ascontiguousarray(centre, dtype=np.float64)
        self.lib.orc_add_sphere(self.h, _ptr(c), float(radius))
        self.info = self.get_info()

    def add_layer(self, k0, k1):
        """Metal on every E component at k in [k0, k1) over all columns."""
        self.lib.orc_add_layer(self.h, int(k0), int(k1))
        self.info = self.get_info()

    def get_materials(self, comp):
        out = np.empty(self.shape, dtype=np.uint8)
        self.lib.orc_get_materials(self.h, comp, _ptr(out, C.c_uint8))
        return out

    def update_h(self):
        self.lib.orc_update_h(self.h)

    def tfsf(self):
        self.lib.orc_tfsf(self.h)

    def update_e(self):
        self.lib.orc_update_e(self.h)

    def end_step(self):
        self.lib.orc_end_step(self.h)

    def step(self, n=1):
        self.lib.orc_step(self.h, int(n))

    def get_field(self, comp):
        out = np.empty(self.shape)
        self.lib.orc_get_field(self.h, comp, _ptr(out))
        return out

    def set_field(self, comp, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.shape)
        self.lib.orc_set_field(self.h, comp, _ptr(a))

    def get_state(self, which, comp):
        """Dispersive state of E component comp: which 0 = hydrodynamic density ND,
        1 = summed CP polarisation, 2 = Drude pole (velocity with the hydrodynamic model)."""
        out = np.empty(self.shape)
        if self.lib.orc_get_state(self.h, int(which), int(comp), _ptr(out)) != 0:
            raise RuntimeError(self.lib.orc_last_error().decode())
        return out

    def get_line(self):
        L = self.info.inc_length
        e, h = np.empty(L), np.empty(L)
        self.lib.orc_get_line(self.h, _ptr(e), _ptr(h))
        return e, h

    def spectrum_setup(self):
        F = self.info.num_freq
        f, ei = np.empty(F), np.empty(2 * F)
        self.lib.orc_get_spectrum_setup(self.h, _ptr(f), _ptr(ei))
        return f, ei[0::2] + 1j * ei[1::2]

    def get_dft(self, face, comp):
        na, nb = face_shape(self.info, face)
        out = np.empty((self.info.num_freq, na, nb, 2))
        self.lib.orc_get_dft(self.h, face, comp, _ptr(out))
        return out[..., 0] + 1j * out[..., 1]

    def flux(self):
        F = self.info.num_freq
        sca, ab, ext, raw = np.empty(F), np.empty(F), np.empty(F), np.empty(12 * F)
        self.lib.orc_flux(self.h, *[_ptr(a) for a in (sca, ab, ext, raw)])
        return {"sca": sca, "abs": ab, "ext": ext, "raw": raw.reshape(12, F)}

    def reflectance(self):
        """Periodic spectra: R, T (physical) and the reference's two columns."""
        F = self.info.num_freq
        out = [np.empty(F) for _ in range(4)]
        if self.lib.orc_reflectance(self.h, *[_ptr(a) for a in out]) != 0:
            raise RuntimeError(self.lib.orc_last_error().decode())
        return dict(zip(("R", "T", "ref_r", "ref_t"), out))

    def coeffs(self):
        out = np.empty(64)
        n = self.lib.orc_coeffs(self.h, _ptr(out))
        return dict(zip(COEFF_NAMES, out[:n]))

    def cpml(self, axis, which):
        n = max(self.shape) + 2
        out = np.empty(n)
        m = self.lib.orc_cpml(self.h, axis, CPML_NAMES.index(which), _ptr(out))
        return out[:m].copy()

    def threads(self):
        return self.lib.orc_threads()

    def set_design(self, lo, rho):
        """Density block at lo with shape rho.shape (i-major, k fastest)."""
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        self.design_shape = rho.shape
        a, b = (C.c_int * 3)(*lo), (C.c_int * 3)(*rho.shape)
        self.lib.orc_set_design(self.h, C.byref(a), C.byref(b), _ptr(rho.reshape(-1)))

    def set_objective(self, r0, freqs):
        f = np.ascontiguousarray(freqs, dtype=np.float64)
        self._obj_freq = f
        a = (C.c_int * 3)(*r0)
        self.lib.orc_set_objective(self.h, C.byref(a), len(f), _ptr(f))

    def objective(self):
        return self.lib.orc_objective(self.h)

    def objective_terms(self):
        """Complex sums sum_t E_c^{t+1}(r0) e^{i 2 pi f (t+1) dt}, shape [nf, 3]."""
        n = len(self._obj_freq)
        out = np.empty(n * 6)
        self.lib.orc_objective_terms(self.h, _ptr(out))
        return (out[0::2] + 1j * out[1::2]).reshape(n, 3)

    def adjoint(self, nsteps):
        """Runs nsteps forward then the reverse sweep; returns (J, dJ/drho)."""
        n = int(np.prod(self.design_shape))
        g = np.empty(n)
        J = C.c_double()
        rc = self.lib.orc_adjoint(self.h, int(nsteps), C.byref(J), _ptr(g))
        if rc != 0:
            raise RuntimeError(self.lib.orc_last_error().decode())
        return J.value, g.reshape(self.design_shape)
