"""Python binding of the MI355X FDTD solver over its C ABI (include/fdtd.h).

The compute path is the HIP library libfdtd_hip.so built in-tree by
`make -C fdtd-cuda_amd`; importing this module never falls back to any other
implementation: a missing library or device raises.
"""
import ctypes as C
import os

import numpy as np

from .abi import (EX, EY, EZ, HX, HY, HZ, FDTD_OK, GridInfo, Params, face_shape)  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libfdtd_hip.so")

EXPORTS = [
    "fdtd_default_params", "fdtd_create", "fdtd_destroy", "fdtd_last_error", "fdtd_get_info",
    "fdtd_set_materials", "fdtd_add_sphere", "fdtd_get_materials", "fdtd_update_h", "fdtd_tfsf", "fdtd_update_e",
    "fdtd_dft_accumulate", "fdtd_step", "fdtd_sync", "fdtd_get_field", "fdtd_set_field",
    "fdtd_get_slice", "fdtd_get_line", "fdtd_get_spectrum_setup", "fdtd_get_dft", "fdtd_flux", "fdtd_kernel_timing",
    "fdtd_set_design", "fdtd_adjoint", "fdtd_add_rod", "fdtd_set_point_monitor", "fdtd_point_spectrum",
    "fdtd_batch_create", "fdtd_batch_run", "fdtd_batch_tile", "fdtd_batch_last_error", "fdtd_batch_destroy",
    "fdtd_checkpoint_size", "fdtd_checkpoint_save", "fdtd_checkpoint_load",
    "fdtd_reflectance", "fdtd_add_layer", "fdtd_flux_noalloc",
]

# largest NUM_freq fdtd_flux_noalloc accepts (its fixed device array)
NOALLOC_MAX_FREQ = 256

BATCH_NO_GRAPH = 1
BATCH_NO_TUNE = 2
BATCH_GENERAL = 4

_lib = None


def load_library(path=LIB_PATH):
    """Loads libfdtd_hip.so and declares the ABI signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library not built: {path} (run `make -C fdtd-cuda_amd`)")
    lib = C.CDLL(path)
    vp, ip, dp, fp, u8p = C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_uint8)
    sig = {
        "fdtd_default_params": (None, [C.POINTER(Params)]),
        "fdtd_create": (ip, [C.POINTER(Params), C.POINTER(vp)]),
        "fdtd_destroy": (ip, [vp]),
        "fdtd_last_error": (C.c_char_p, [vp]),
        "fdtd_get_info": (ip, [vp, C.POINTER(GridInfo)]),
        "fdtd_set_materials": (ip, [vp, u8p, u8p, u8p]),
        "fdtd_add_sphere": (ip, [vp, C.POINTER(C.c_double * 3), C.c_double]),
        "fdtd_get_materials": (ip, [vp, ip, u8p]),
        "fdtd_update_h": (ip, [vp]),
        "fdtd_tfsf": (ip, [vp]),
        "fdtd_update_e": (ip, [vp]),
        "fdtd_dft_accumulate": (ip, [vp]),
        "fdtd_step": (ip, [vp, ip]),
        "fdtd_sync": (ip, [vp]),
        "fdtd_get_field": (ip, [vp, ip, fp]),
        "fdtd_set_field": (ip, [vp, ip, fp]),
        "fdtd_get_line": (ip, [vp, dp, dp]),
        "fdtd_get_spectrum_setup": (ip, [vp, dp, dp]),
        "fdtd_get_slice": (ip, [vp, ip, ip, ip, C.POINTER(C.c_float)]),
        "fdtd_get_dft": (ip, [vp, ip, ip, dp]),
        "fdtd_flux": (ip, [vp, dp, dp, dp, dp]),
        "fdtd_flux_noalloc": (ip, [vp, dp, dp, dp, dp]),
        "fdtd_kernel_timing": (ip, [vp, ip, dp, C.POINTER(C.c_long)]),
        "fdtd_set_design": (ip, [vp, C.POINTER(C.c_int * 3), C.POINTER(C.c_int * 3), fp]),
        "fdtd_adjoint": (ip, [vp, ip, ip, C.POINTER(C.c_int * 3), ip, dp, dp, dp]),
        "fdtd_add_rod": (ip, [vp, C.POINTER(C.c_double * 3), ip, C.c_double, C.c_double]),
        "fdtd_set_point_monitor": (ip, [vp, C.POINTER(C.c_int * 3), ip]),
        "fdtd_point_spectrum": (ip, [vp, ip, dp, dp, dp]),
        "fdtd_batch_create": (ip, [C.POINTER(vp), ip, ip, C.POINTER(vp)]),
        "fdtd_batch_run": (ip, [vp, ip]),
        "fdtd_batch_tile": (ip, [vp, ip, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "fdtd_batch_last_error": (C.c_char_p, [vp]),
        "fdtd_batch_destroy": (ip, [vp]),
        "fdtd_checkpoint_size": (ip, [vp, C.POINTER(C.c_size_t)]),
        "fdtd_checkpoint_save": (ip, [vp, vp, C.c_size_t]),
        "fdtd_checkpoint_load": (ip, [vp, vp, C.c_size_t]),
        "fdtd_reflectance": (ip, [vp, dp, dp, dp, dp]),
        "fdtd_add_layer": (ip, [vp, ip, ip]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def default_params():
    p = Params()
    load_library().fdtd_default_params(C.byref(p))
    return p


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class FDTDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fdtd error {code}: {msg}")
        self.code = code


class Simulation:
    """One solver handle (FP32 state on the GPU)."""

    def __init__(self, params):
        self.lib = load_library()
        self.h = C.c_void_p()
        rc = self.lib.fdtd_create(C.byref(params), C.byref(self.h))
        if rc != FDTD_OK:
            raise FDTDError(rc, self.lib.fdtd_last_error(None).decode())
        self.params = params
        self.info = self.get_info()
        self.shape = (self.info.NX, self.info.NY, self.info.NZ)

    def _check(self, rc):
        if rc != FDTD_OK:
            raise FDTDError(rc, self.lib.fdtd_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.lib.fdtd_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_info(self):
        g = GridInfo()
        self._check(self.lib.fdtd_get_info(self.h, C.byref(g)))
        return g

    def set_materials(self, mx, my, mz):
        arrs = [np.ascontiguousarray(m, dtype=np.uint8).reshape(-1) for m in (mx, my, mz)]
        self._check(self.lib.fdtd_set_materials(self.h, *[_ptr(a, C.c_uint8) for a in arrs]))
        self.info = self.get_info()

    def add_sphere(self, centre, radius):
        c = (C.c_double * 3)(*[float(x) for x in centre])
        self._check(self.lib.fdtd_add_sphere(self.h, C.byref(c), float(radius)))
        self.info = self.get_info()

    def add_rod(self, centre, axis, length, radius):
        """Nanorod with hemispherical caps (length, radius in metres) along axis 0..2."""
        c = (C.c_double * 3)(*[float(x) for x in centre])
        self._check(self.lib.fdtd_add_rod(self.h, C.byref(c), int(axis), float(length), float(radius)))
        self.info = self.get_info()

    def set_point_monitor(self, r, capacity):
        """Records E at cell r after each of the next `capacity` steps."""
        a = (C.c_int * 3)(*[int(x) for x in r])
        self._check(self.lib.fdtd_set_point_monitor(self.h, C.byref(a), int(capacity)))

    def point_spectrum(self, freqs):
        """(|E(f)|^2 per frequency, complex amplitudes [nf, 3]) of the point monitor."""
        f = np.ascontiguousarray(freqs, dtype=np.float64)
        e2 = np.empty(len(f))
        amp = np.empty(len(f) * 6)
        self._check(self.lib.fdtd_point_spectrum(self.h, len(f), _ptr(f, C.c_double), _ptr(e2, C.c_double),
                                                 _ptr(amp, C.c_double)))
        return e2, (amp[0::2] + 1j * amp[1::2]).reshape(len(f), 3)

    def set_design(self, lo, rho):
        """Design block of densities rho (shape n, i-major, k fastest) at cell lo."""
        r = np.ascontiguousarray(rho, dtype=np.float32)
        self.design_shape = r.shape
        a, b = (C.c_int * 3)(*[int(x) for x in lo]), (C.c_int * 3)(*r.shape)
        self._check(self.lib.fdtd_set_design(self.h, C.byref(a), C.byref(b), _ptr(r.reshape(-1), C.c_float)))
        self.info = self.get_info()

    def adjoint(self, nsteps, r0, freqs, checkpoint=64):
        """nsteps forward steps, then the discrete adjoint sweep of the point
        objective J = sum_f sum_c |sum_t E_c^{t+1}(r0) exp(i 2 pi f (t+1) dt)|^2.
        checkpoint > 0: checkpoints every `checkpoint` steps and segment
        recompute; 0: the design-box trajectory is kept instead.
        Returns (J, dJ/drho over the design block)."""
        f = np.ascontiguousarray(freqs, dtype=np.float64)
        g = np.empty(int(np.prod(self.design_shape)))
        J = C.c_double()
        r = (C.c_int * 3)(*[int(x) for x in r0])
        self._check(self.lib.fdtd_adjoint(self.h, int(nsteps), int(checkpoint), C.byref(r), len(f),
                                          _ptr(f, C.c_double), C.byref(J), _ptr(g, C.c_double)))
        return J.value, g.reshape(self.design_shape)

    def get_materials(self, comp):
        out = np.empty(self.shape, dtype=np.uint8)
        self._check(self.lib.fdtd_get_materials(self.h, comp, _ptr(out, C.c_uint8)))
        return out

    def update_h(self):
        self._check(self.lib.fdtd_update_h(self.h))

    def tfsf(self):
        self._check(self.lib.fdtd_tfsf(self.h))

    def update_e(self):
        self._check(self.lib.fdtd_update_e(self.h))

    def dft_accumulate(self):
        self._check(self.lib.fdtd_dft_accumulate(self.h))

    def split_step(self, n=1):
        for _ in range(n):
            self.update_h()
            self.tfsf()
            self.update_e()
            self.dft_accumulate()

    def step(self, n=1):
        self._check(self.lib.fdtd_step(self.h, int(n)))

    def sync(self):
        self._check(self.lib.fdtd_sync(self.h))

    def get_field(self, comp):
        out = np.empty(self.shape, dtype=np.float32)
        self._check(self.lib.fdtd_get_field(self.h, comp, _ptr(out, C.c_float)))
        return out

    def set_field(self, comp, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32).reshape(self.shape)
        self._check(self.lib.fdtd_set_field(self.h, comp, _ptr(a, C.c_float)))

    def get_line(self):
        L = self.info.inc_length
        e, h = np.empty(L), np.empty(L)
        self._check(self.lib.fdtd_get_line(self.h, _ptr(e, C.c_double), _ptr(h, C.c_double)))
        return e, h

    def get_slice(self, comp, axis, index):
        """One plane of a field component (0-5) or, with the hydrodynamic model,
        of the density perturbation (6): axis 0 -> [j, k], 1 -> [i, k], 2 -> [i, j]."""
        i = self.info
        shape = {0: (i.NY, i.NZ), 1: (i.NX, i.NZ), 2: (i.NX, i.NY)}[axis]
        out = np.empty(shape, dtype=np.float32)
        self._check(self.lib.fdtd_get_slice(self.h, int(comp), int(axis), int(index), _ptr(out, C.c_float)))
        return out

    def spectrum_setup(self):
        F = self.info.num_freq
        f, ei = np.empty(F), np.empty(2 * F)
        self._check(self.lib.fdtd_get_spectrum_setup(self.h, _ptr(f, C.c_double), _ptr(ei, C.c_double)))
        return f, ei[0::2] + 1j * ei[1::2]

    def get_dft(self, face, comp):
        na, nb = face_shape(self.info, face)
        out = np.empty((self.info.num_freq, na, nb, 2))
        self._check(self.lib.fdtd_get_dft(self.h, face, comp, _ptr(out, C.c_double)))
        return out[..., 0] + 1j * out[..., 1]

    def flux(self):
        F = self.info.num_freq
        sca, ab, ext, raw = np.empty(F), np.empty(F), np.empty(F), np.empty(12 * F)
        self._check(self.lib.fdtd_flux(self.h, *[_ptr(a, C.c_double) for a in (sca, ab, ext, raw)]))
        return {"sca": sca, "abs": ab, "ext": ext, "raw": raw.reshape(12, F)}

    def flux_noalloc(self):
        """flux() through fdtd_flux_noalloc (no device allocation; the same bits)."""
        F = self.info.num_freq
        sca, ab, ext, raw = np.empty(F), np.empty(F), np.empty(F), np.empty(12 * F)
        self._check(self.lib.fdtd_flux_noalloc(self.h, *[_ptr(a, C.c_double) for a in (sca, ab, ext, raw)]))
        return {"sca": sca, "abs": ab, "ext": ext, "raw": raw.reshape(12, F)}

    def add_layer(self, k0, k1):
        """Metal on every E component at k in [k0, k1) over the whole x-y extent (periodic runs)."""
        self._check(self.lib.fdtd_add_layer(self.h, int(k0), int(k1)))
        self.info = self.get_info()

    def reflectance(self):
        """Periodic spectra: R, T (physically normalised) and the reference's columns ref_r, ref_t."""
        F = self.info.num_freq
        out = [np.empty(F) for _ in range(4)]
        self._check(self.lib.fdtd_reflectance(self.h, *[_ptr(a, C.c_double) for a in out]))
        return dict(zip(("R", "T", "ref_r", "ref_t"), out))

    def checkpoint(self):
        """The whole simulation state as a bytes-like numpy array (fdtd_checkpoint_save)."""
        n = C.c_size_t()
        self._check(self.lib.fdtd_checkpoint_size(self.h, C.byref(n)))
        buf = np.empty(n.value, dtype=np.uint8)
        self._check(self.lib.fdtd_checkpoint_save(self.h, buf.ctypes.data_as(C.c_void_p), n.value))
        return buf

    def restore(self, buf):
        """Loads an image written by checkpoint() (fdtd_checkpoint_load)."""
        b = np.ascontiguousarray(buf, dtype=np.uint8)
        self._check(self.lib.fdtd_checkpoint_load(self.h, b.ctypes.data_as(C.c_void_p), b.size))
        self.info = self.get_info()

    def kernel_timing(self, enable):
        ms, n = C.c_double(), C.c_long()
        self._check(self.lib.fdtd_kernel_timing(self.h, int(enable), C.byref(ms), C.byref(n)))
        return ms.value, n.value


class Batch:
    """Many simulations advanced together (fdtd_batch_*): one launch per
    kernel per step for all of them, a HIP graph per sampling period, tiles
    picked by the autotuner.  Keeps references to the simulations."""

    def __init__(self, sims, flags=0):
        self.lib = load_library()
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[s.h.value for s in self.sims])
        self.h = C.c_void_p()
        rc = self.lib.fdtd_batch_create(arr, len(self.sims), int(flags), C.byref(self.h))
        if rc != FDTD_OK:
            raise FDTDError(rc, self.lib.fdtd_batch_last_error(None).decode())

    def _check(self, rc):
        if rc != FDTD_OK:
            raise FDTDError(rc, self.lib.fdtd_batch_last_error(self.h).decode())

    def run(self, nsteps):
        self._check(self.lib.fdtd_batch_run(self.h, int(nsteps)))

    def tile(self, i):
        """(path, a, c): path 1 row-blocked kernel (rows per thread, x chunk), 0 general (block rows, x chunk)."""
        path, a, c = C.c_int(), C.c_int(), C.c_int()
        self._check(self.lib.fdtd_batch_tile(self.h, int(i), C.byref(path), C.byref(a), C.byref(c)))
        return path.value, a.value, c.value

    def close(self):
        if self.h:
            self.lib.fdtd_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
