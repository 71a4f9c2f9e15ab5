"""TEST INFRASTRUCTURE — NOT PART OF THE PRODUCT.

ctypes wrapper over oracle/liboracle.so, the serial/OpenMP C++ transcription
of the reference kernels (src/lbm.cu, src/init.cu, src/main.cu of
kreeyto/MULTIC-BUBBLE-CUDA). Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg use it, as the checker or the CPU baseline.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

PERIODIC, REFERENCE = 0, 1
W_LITERAL, W_UNIT_SUM = 0, 1

# Field ids of ora_field_ptr (the reference's buffers, var.cu:13-18).
FIELDS = {
    "f": 0, "g": 1, "g_out": 2, "phi": 3, "normx": 4, "normy": 5, "normz": 6,
    "indicator": 7, "curvature": 8, "ffx": 9, "ffy": 10, "ffz": 11,
    "ux": 12, "uy": 13, "uz": 14, "rho": 15, "pxx": 16, "pyy": 17, "pzz": 18,
    "pxy": 19, "pxz": 20, "pyz": 21,
}
# Stage ids of ora_stage, named after the reference kernels.
STAGES = {
    "phiCalc": 0, "gradCalc": 1, "curvatureCalc": 2, "momentiCalc": 3,
    "collisionFluid": 4, "collisionPhase": 5, "streamingCalc": 6,
    "fgBoundary": 7, "boundaryConditions": 8, "movingWallBC": 9,
    "initTensor": 10, "initDist": 11,
}


class _Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("bc_mode", ctypes.c_int), ("weight_set", ctypes.c_int),
        ("moving_wall", ctypes.c_int),
        ("tau_f", ctypes.c_double), ("tau_g", ctypes.c_double),
        ("sigma", ctypes.c_double), ("sharp_c", ctypes.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(_LIB_PATH)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.ora_create.restype = vp
        L.ora_create.argtypes = [i, ctypes.POINTER(_Params)]
        L.ora_destroy.argtypes = [vp]
        L.ora_init.argtypes = [vp, i, ctypes.POINTER(d), d]
        L.ora_step.argtypes = [vp, i]
        L.ora_macros.argtypes = [vp]
        L.ora_stage.argtypes = [vp, i]
        L.ora_field_ptr.restype = vp
        L.ora_field_ptr.argtypes = [vp, i]
        L.ora_field_len.restype = ctypes.c_size_t
        L.ora_field_len.argtypes = [vp, i]
        L.ora_sum_phi.restype = d
        L.ora_sum_phi.argtypes = [vp]
        L.ora_kinetic.restype = d
        L.ora_kinetic.argtypes = [vp]
        L.ora_weights.argtypes = [vp, ctypes.POINTER(d)]
        L.ora_set_threads.argtypes = [i]
        L.ora_max_threads.restype = i
        L.ora_time_steps.restype = d
        L.ora_time_steps.argtypes = [vp, i]
        L.ora_tanh_scan.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_longlong)]
        _lib = L
    return _lib


class Oracle:
    """One lattice state of the transcription, FP32 (prec=32) or FP64."""

    def __init__(self, nx, ny, nz, *, prec=32, bc_mode=PERIODIC, weights=W_UNIT_SUM,
                 tau_f=0.505, tau_g=1.0, sigma=0.1, sharp_c=-1.0, moving_wall=False):
        self.shape = (nz, ny, nx)
        self.prec = prec
        self.dtype = np.float64 if prec == 64 else np.float32
        p = _Params(nx, ny, nz, bc_mode, weights, int(moving_wall),
                    float(tau_f), float(tau_g), float(sigma), float(sharp_c))
        self._h = lib().ora_create(prec, ctypes.byref(p))
        if not self._h:
            raise ValueError("oracle: invalid lattice parameters")

    def close(self):
        if self._h:
            lib().ora_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self, bubbles, width=3.0):
        """bubbles: iterable of (cx, cy, cz, R); the tanh profile of init.cu."""
        b = np.ascontiguousarray(np.asarray(bubbles, dtype=np.float64).reshape(-1, 4))
        lib().ora_init(self._h, b.shape[0], b.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                       float(width))

    def step(self, n=1):
        lib().ora_step(self._h, int(n))

    def macros(self):
        lib().ora_macros(self._h)

    def stage(self, name):
        lib().ora_stage(self._h, STAGES[name])

    def field(self, name):
        """Zero-copy numpy view, (nz, ny, nx) or (19, nz, ny, nx)."""
        fid = FIELDS[name]
        n = lib().ora_field_len(self._h, fid)
        ptr = lib().ora_field_ptr(self._h, fid)
        ctype = ctypes.c_double if self.prec == 64 else ctypes.c_float
        arr = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), shape=(n,))
        nz, ny, nx = self.shape
        if n == 19 * nx * ny * nz:
            return arr.reshape(19, nz, ny, nx)
        return arr.reshape(nz, ny, nx)

    def sum_phi(self):
        return lib().ora_sum_phi(self._h)

    def kinetic(self):
        return lib().ora_kinetic(self._h)

    def weights(self):
        w = (ctypes.c_double * 19)()
        lib().ora_weights(self._h, w)
        return np.array(w[:])

    def time_steps(self, n):
        return lib().ora_time_steps(self._h, int(n))


def set_threads(n):
    lib().ora_set_threads(int(n))


def tanh_scan(lo_bits, hi_bits):
    """(compared, differing, more than 1 ulp apart): the FP32 tanh of the
    initial profile against the C library's double tanh rounded to float, for
    every float bit pattern in [lo_bits, hi_bits] (non-negative floats)."""
    out = (ctypes.c_longlong * 3)()
    lib().ora_tanh_scan(int(lo_bits), int(hi_bits), out)
    return out[0], out[1], out[2]


def max_threads():
    return lib().ora_max_threads()
