"""Python binding of libmcb (include/mcb.h, include/mcb_ops.h) over ctypes.

libmcb is the MI355X-native D3Q19 phase-field bubble LBM: HIP kernels for
gfx950 behind a C ABI. This module only marshals arguments; every step runs
in libmcb on the GPU. If libmcb.so is missing, importing works but any call
raises MissingLibrary: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# MCB_LIBRARY selects another build of libmcb (e.g. a variant for an A/B timing).
LIB_PATH = os.environ.get("MCB_LIBRARY") or os.path.join(_PKG_DIR, "libmcb.so")
DRIVER_PATH = os.path.join(_PKG_DIR, "bin", "mcb")

# mcb_status
OK, ERR_INVALID_ARG, ERR_NO_DEVICE, ERR_OUT_OF_MEMORY, ERR_HIP, ERR_UNSUPPORTED, ERR_IO, ERR_NONFINITE, \
    ERR_CAPACITY = range(9)
# enums
BC_PERIODIC, BC_REFERENCE = 0, 1
WEIGHTS_LITERAL, WEIGHTS_UNIT_SUM = 0, 1
STORE_FP32, STORE_FP16, STORE_BF16 = 0, 1, 2
PATH_AUTO, PATH_FUSED, PATH_STAGED = 0, 1, 2
NORMALS_AUTO, NORMALS_STORED, NORMALS_RECOMPUTED = 0, 1, 2
F_POPULATIONS, F_MOMENTS = 0, 1
FIELDS = {"phi": 0, "rho": 1, "ux": 2, "uy": 3, "uz": 4, "fx": 5, "fy": 6, "fz": 7,
          "normx": 8, "normy": 9, "normz": 10, "f": 16, "g": 17}
MAX_BUBBLES = 8
OPS_MAX_BUBBLES = 8


class MissingLibrary(RuntimeError):
    pass


class McbError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{msg} (status {status})")
        self.status = status


class Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("bc_mode", ctypes.c_int), ("weight_set", ctypes.c_int), ("storage", ctypes.c_int),
        ("path", ctypes.c_int), ("device", ctypes.c_int),
        ("tau_f", ctypes.c_float), ("tau_g", ctypes.c_float), ("sigma", ctypes.c_float),
        ("sharp_c", ctypes.c_float), ("moving_wall", ctypes.c_int),
        ("normals", ctypes.c_int), ("f_form", ctypes.c_int), ("reserved", ctypes.c_int * 5),
    ]


class Bubble(ctypes.Structure):
    _fields_ = [("cx", ctypes.c_float), ("cy", ctypes.c_float), ("cz", ctypes.c_float), ("r", ctypes.c_float)]


class Case(ctypes.Structure):
    _fields_ = [
        ("params", Params), ("n_bubbles", ctypes.c_int), ("bubbles", Bubble * MAX_BUBBLES),
        ("width", ctypes.c_float), ("nsteps", ctypes.c_int), ("save_every", ctypes.c_int),
    ]


class Request(ctypes.Structure):
    _fields_ = [("c", Case), ("arrival_s", ctypes.c_double), ("phi_out", ctypes.POINTER(ctypes.c_float))]


class RequestResult(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int), ("reserved", ctypes.c_int),
        ("arrival_s", ctypes.c_double), ("admit_s", ctypes.c_double), ("finish_s", ctypes.c_double),
        ("sum_phi", ctypes.c_double), ("kinetic", ctypes.c_double), ("lattice_updates", ctypes.c_longlong),
    ]


class ServeConfig(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int), ("storage", ctypes.c_int), ("hbm_budget", ctypes.c_size_t),
        ("max_cases", ctypes.c_int), ("tick_steps", ctypes.c_int), ("sequential", ctypes.c_int),
        ("reserved", ctypes.c_int * 7),
    ]


class ServeStats(ctypes.Structure):
    _fields_ = [
        ("makespan_s", ctypes.c_double), ("lattice_updates", ctypes.c_double), ("mlups", ctypes.c_double),
        ("p50_s", ctypes.c_double), ("p99_s", ctypes.c_double), ("mean_s", ctypes.c_double),
        ("max_s", ctypes.c_double), ("completed", ctypes.c_int), ("rejected", ctypes.c_int),
        ("peak_cases", ctypes.c_int), ("rounds", ctypes.c_int), ("peak_bytes", ctypes.c_double),
        ("admit_s", ctypes.c_double), ("enqueue_s", ctypes.c_double), ("wait_s", ctypes.c_double),
        ("retire_s", ctypes.c_double), ("idle_s", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class OpConsts(ctypes.Structure):
    _fields_ = [
        ("w", ctypes.c_float * 3), ("omega", ctypes.c_float), ("sharp_c", ctypes.c_float),
        ("sigma", ctypes.c_float), ("omega_g", ctypes.c_float), ("tau_g_is_one", ctypes.c_int),
        ("bc_mode", ctypes.c_int), ("n_bubbles", ctypes.c_int),
        ("bubbles", ctypes.c_float * (4 * OPS_MAX_BUBBLES)), ("width", ctypes.c_float),
    ]


# Exported symbols and their prototypes (restype, argtypes).
_vp, _i, _sz, _f, _d, _cp = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, ctypes.c_float, ctypes.c_double, \
    ctypes.c_char_p
_PP = ctypes.POINTER(Params)
_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_OCP = ctypes.POINTER(OpConsts)
PROTOTYPES = {
    "mcb_params_reference": (_i, [_PP]),
    "mcb_validate_params": (_i, [_PP]),
    "mcb_footprint_bytes": (_sz, [_PP]),
    "mcb_create": (_i, [_PP, _vp, ctypes.POINTER(_vp)]),
    "mcb_destroy": (_i, [_vp]),
    "mcb_init_bubbles": (_i, [_vp, _i, ctypes.POINTER(Bubble), _f]),
    "mcb_init_phi": (_i, [_vp, _fp, _sz]),
    "mcb_step": (_i, [_vp, _i]),
    "mcb_step_timed": (_i, [_vp, _i, _dp]),
    "mcb_compute_macros": (_i, [_vp]),
    "mcb_read_field": (_i, [_vp, _i, _fp, _sz]),
    "mcb_write_field": (_i, [_vp, _i, _fp, _sz]),
    "mcb_reduce": (_i, [_vp, _dp, _dp]),
    "mcb_synchronize": (_i, [_vp]),
    "mcb_step_kernel_times": (_i, [_vp, _i, _dp, _i]),
    "mcb_set_tiling": (_i, [_vp, _i, _i, _i]),
    "mcb_get_tiling": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "mcb_tiling_count": (_i, []),
    "mcb_tiling_at": (_i, [_i, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "mcb_abi_version": (_i, []),
    "mcb_last_error": (_cp, [_vp]),
    "mcb_status_string": (_cp, [_i]),
    "mcb_get_params": (_i, [_vp, _PP]),
    "mcb_case_preset": (_i, [_cp, ctypes.c_uint64, ctypes.POINTER(Case)]),
    "mcb_host_init_phi": (_i, [_PP, _i, ctypes.POINTER(Bubble), _f, _fp, _sz]),
    "mcb_write_info_file": (_i, [_cp, _cp, _cp, _i, _i, _i, _i, _i, _f, _d]),
    "mcb_write_field_file": (_i, [_cp, _cp, _cp, _i, _fp, _sz]),
    "mcb_device_count": (_i, [ctypes.POINTER(_i)]),
    "mcb_device_alloc": (_i, [_i, _sz, ctypes.POINTER(_vp)]),
    "mcb_device_free": (_i, [_vp]),
    "mcb_memcpy_htod": (_i, [_vp, _vp, _sz]),
    "mcb_memcpy_dtoh": (_i, [_vp, _vp, _sz]),
    "mcb_device_synchronize": (_i, []),
    "mcb_op_consts_from_params": (_i, [_PP, _OCP]),
    "mcb_bench_copy": (_i, [_i, _sz, _i, _dp]),
    "mcb_autotune": (_i, [_PP, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i), _dp]),
    "mcb_case_bytes": (_sz, [_PP]),
    "mcb_batch_create": (_i, [_i, _i, _vp, ctypes.POINTER(_vp)]),
    "mcb_batch_destroy": (_i, [_vp]),
    "mcb_batch_add": (_i, [_vp, _PP, _i, ctypes.POINTER(Bubble), _f, ctypes.POINTER(_i)]),
    "mcb_batch_add_phi": (_i, [_vp, _PP, _fp, _sz, ctypes.POINTER(_i)]),
    "mcb_batch_remove": (_i, [_vp, _i]),
    "mcb_batch_step": (_i, [_vp, _i]),
    "mcb_batch_step_timed": (_i, [_vp, _i, _dp]),
    "mcb_batch_compute_macros": (_i, [_vp, _i]),
    "mcb_batch_read_field": (_i, [_vp, _i, _i, _fp, _sz]),
    "mcb_batch_write_field": (_i, [_vp, _i, _i, _fp, _sz]),
    "mcb_batch_reduce": (_i, [_vp, _i, _dp, _dp]),
    "mcb_batch_case_steps": (_i, [_vp, _i, ctypes.POINTER(ctypes.c_longlong)]),
    "mcb_batch_set_tiling": (_i, [_vp, _i, _i, _i, _i]),
    "mcb_batch_synchronize": (_i, [_vp]),
    "mcb_batch_count": (_i, [_vp]),
    "mcb_batch_bytes": (_sz, [_vp]),
    "mcb_batch_last_error": (_cp, [_vp]),
    "mcb_state_size": (_i, [_vp, ctypes.POINTER(_sz)]),
    "mcb_save_state": (_i, [_vp, _vp, _sz]),
    "mcb_load_state": (_i, [_vp, _vp, _sz]),
    "mcb_state_info": (_i, [_vp, _sz, _PP, ctypes.POINTER(ctypes.c_longlong)]),
    "mcb_batch_state_size": (_i, [_vp, _i, ctypes.POINTER(_sz)]),
    "mcb_batch_save_state": (_i, [_vp, _i, _vp, _sz]),
    "mcb_batch_load_state": (_i, [_vp, _i, _vp, _sz]),
    "mcb_batch_add_state": (_i, [_vp, _vp, _sz, ctypes.POINTER(_i)]),
    "mcb_serve_run": (_i, [ctypes.POINTER(ServeConfig), ctypes.POINTER(Request), _i, ctypes.POINTER(RequestResult),
                           ctypes.POINTER(ServeStats)]),
    "mcb_serve_trace": (_i, [ctypes.c_uint64, _i, _d, _d, ctypes.POINTER(Request)]),
    "mcb_reduce_fields_noalloc": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _vp, _vp]),
}
# mcb_ops.h: reference kernel names -> number of device pointer arguments.
OP_POINTERS = {
    "initTensor": 7, "initPhase": 1, "initDist": 4, "phiCalc": 2, "gradCalc": 5, "curvatureCalc": 8,
    "momentiCalc": 14, "collisionFluid": 14, "collisionPhase": 8, "streamingCalc": 2, "fgBoundary": 4,
    "boundaryConditions": 1, "movingWallBC": 2,
}
for _name, _np in OP_POINTERS.items():
    PROTOTYPES["mcb_" + _name] = (_i, [_vp] * _np + [_i, _i, _i, _OCP, _vp])

_lib = None


def lib():
    """The loaded libmcb (raises MissingLibrary if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MissingLibrary(f"libmcb.so not built at {LIB_PATH}; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def status_string(st):
    return lib().mcb_status_string(st).decode()


def check(st, ctx=None, what="libmcb call"):
    if st != OK:
        msg = status_string(st)
        if ctx:
            detail = lib().mcb_last_error(ctx)
            if detail:
                msg += ": " + detail.decode()
        raise McbError(st, f"{what}: {msg}")


def device_count():
    n = ctypes.c_int(0)
    lib().mcb_device_count(ctypes.byref(n))
    return n.value


def bench_copy(nbytes, iters=20, device=0):
    """Milliseconds per plain device copy of nbytes (read nbytes + write nbytes)."""
    ms = ctypes.c_double(0)
    check(lib().mcb_bench_copy(device, nbytes, iters, ctypes.byref(ms)), what="bench_copy")
    return ms.value


def reference_params():
    p = Params()
    check(lib().mcb_params_reference(ctypes.byref(p)))
    return p


def make_params(nx, ny, nz, *, bc_mode=BC_PERIODIC, weights=WEIGHTS_UNIT_SUM, storage=STORE_FP32,
                path=PATH_AUTO, device=0, tau_f=0.505, tau_g=1.0, sigma=0.1, sharp_c=-1.0, moving_wall=False,
                normals=0, f_form=0):
    p = reference_params()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.bc_mode, p.weight_set, p.storage, p.path, p.device = bc_mode, weights, storage, path, device
    p.tau_f, p.tau_g, p.sigma, p.sharp_c = tau_f, tau_g, sigma, sharp_c
    p.moving_wall = int(moving_wall)
    p.normals = int(normals)
    p.f_form = int(f_form)
    return p


def case_preset(name, seed=0):
    c = Case()
    check(lib().mcb_case_preset(name.encode(), ctypes.c_uint64(seed), ctypes.byref(c)), what=f"preset {name}")
    return c


def case_bubbles(c):
    return [(b.cx, b.cy, b.cz, b.r) for b in c.bubbles[:c.n_bubbles]]


def _bubble_array(bubbles):
    bubbles = list(bubbles)
    arr = (Bubble * len(bubbles))()
    for i, (cx, cy, cz, r) in enumerate(bubbles):
        arr[i] = Bubble(cx, cy, cz, r)
    return arr


def host_init_phi(params, bubbles, width=3.0):
    n = params.nx * params.ny * params.nz
    out = np.empty(n, dtype=np.float32)
    arr = _bubble_array(bubbles)
    check(lib().mcb_host_init_phi(ctypes.byref(params), len(arr), arr, width,
                                  out.ctypes.data_as(_fp), n), what="host_init_phi")
    return out.reshape(params.nz, params.ny, params.nx)


def _fptr(a):
    return a.ctypes.data_as(_fp)


class Lattice:
    """One libmcb context: a lattice resident on the GPU."""

    def __init__(self, params, stream=None):
        self.params = params
        self.shape = (params.nz, params.ny, params.nx)
        self.cells = params.nx * params.ny * params.nz
        h = ctypes.c_void_p()
        check(lib().mcb_create(ctypes.byref(params), stream, ctypes.byref(h)), what="mcb_create")
        self._h = h

    @classmethod
    def from_case(cls, case, **overrides):
        # A copy: the overrides apply to this lattice, not to the case.
        p = Params.from_buffer_copy(case.params)
        for k, v in overrides.items():
            setattr(p, k, v)
        lat = cls(p)
        lat.init_bubbles(case_bubbles(case), case.width)
        return lat

    def close(self):
        if getattr(self, "_h", None):
            lib().mcb_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st, what):
        check(st, self._h, what)

    def init_bubbles(self, bubbles, width=3.0):
        arr = _bubble_array(bubbles)
        self._chk(lib().mcb_init_bubbles(self._h, len(arr), arr, width), "mcb_init_bubbles")

    def init_phi(self, phi):
        a = np.ascontiguousarray(phi, dtype=np.float32).ravel()
        self._chk(lib().mcb_init_phi(self._h, _fptr(a), a.size), "mcb_init_phi")

    def step(self, n=1):
        self._chk(lib().mcb_step(self._h, int(n)), "mcb_step")

    def step_timed(self, n):
        ms = ctypes.c_double(0)
        self._chk(lib().mcb_step_timed(self._h, int(n), ctypes.byref(ms)), "mcb_step_timed")
        return ms.value

    def kernel_times(self, n):
        out = (ctypes.c_double * 2)()
        self._chk(lib().mcb_step_kernel_times(self._h, int(n), out, 2), "mcb_step_kernel_times")
        return out[0], out[1]

    def compute_macros(self):
        self._chk(lib().mcb_compute_macros(self._h), "mcb_compute_macros")

    def read(self, name):
        fid = FIELDS[name]
        if fid >= 16:
            out = np.empty(19 * self.cells, dtype=np.float32)
            self._chk(lib().mcb_read_field(self._h, fid, _fptr(out), out.size), f"read {name}")
            return out.reshape((19,) + self.shape)
        out = np.empty(self.cells, dtype=np.float32)
        self._chk(lib().mcb_read_field(self._h, fid, _fptr(out), out.size), f"read {name}")
        return out.reshape(self.shape)

    def write(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32).ravel()
        self._chk(lib().mcb_write_field(self._h, FIELDS[name], _fptr(a), a.size), f"write {name}")

    def reduce(self):
        s, k = ctypes.c_double(0), ctypes.c_double(0)
        st = lib().mcb_reduce(self._h, ctypes.byref(s), ctypes.byref(k))
        if st not in (OK, ERR_NONFINITE):
            self._chk(st, "mcb_reduce")
        return s.value, k.value, st == OK

    def synchronize(self):
        self._chk(lib().mcb_synchronize(self._h), "mcb_synchronize")

    def set_tiling(self, tx, ty, zc):
        self._chk(lib().mcb_set_tiling(self._h, tx, ty, zc), "mcb_set_tiling")

    def tiling(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._chk(lib().mcb_get_tiling(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)), "get_tiling")
        return a.value, b.value, c.value

    def observables(self):
        """phi, rho, ux, uy, uz of the current state (a17: the macros of the next iteration)."""
        self.compute_macros()
        return {k: self.read(k) for k in ("phi", "rho", "ux", "uy", "uz")}

    def save_state(self):
        """The complete lattice state as bytes (mcb_save_state): loading it into
        a Lattice with the same parameters continues the run bit for bit."""
        n = ctypes.c_size_t()
        self._chk(lib().mcb_state_size(self._h, ctypes.byref(n)), "mcb_state_size")
        buf = ctypes.create_string_buffer(n.value)
        self._chk(lib().mcb_save_state(self._h, buf, n.value), "mcb_save_state")
        return buf.raw

    def load_state(self, image):
        buf = _image_buffer(image)
        self._chk(lib().mcb_load_state(self._h, buf, len(buf)), "mcb_load_state")


def _image_buffer(image):
    b = bytes(image)
    return (ctypes.c_char * len(b)).from_buffer_copy(b)


def state_info(image):
    """(Params, steps) recorded in a state image."""
    buf = _image_buffer(image)
    p, steps = Params(), ctypes.c_longlong()
    check(lib().mcb_state_info(buf, len(buf), ctypes.byref(p), ctypes.byref(steps)), what="mcb_state_info")
    return p, steps.value


def autotune(params):
    """Time every tile shape on a scratch lattice; returns (tx, ty, zc, ms_per_step)."""
    a, b, c, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    check(lib().mcb_autotune(ctypes.byref(params), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                             ctypes.byref(ms)), what="mcb_autotune")
    return a.value, b.value, c.value, ms.value


def case_bytes(params):
    return lib().mcb_case_bytes(ctypes.byref(params))


class Batch:
    """Many independent periodic cases advanced together (mcb_batch_*)."""

    def __init__(self, storage=STORE_FP32, device=0, stream=None):
        h = ctypes.c_void_p()
        check(lib().mcb_batch_create(device, storage, stream, ctypes.byref(h)), what="mcb_batch_create")
        self._h = h
        self.storage = storage
        self.shapes = {}

    def _chk(self, st, what):
        if st != OK:
            detail = lib().mcb_batch_last_error(self._h)
            raise McbError(st, f"{what}: {status_string(st)}" + (f": {detail.decode()}" if detail else ""))

    def close(self):
        if getattr(self, "_h", None):
            lib().mcb_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, params, bubbles, width=3.0):
        arr = _bubble_array(bubbles)
        cid = ctypes.c_int()
        self._chk(lib().mcb_batch_add(self._h, ctypes.byref(params), len(arr), arr, width, ctypes.byref(cid)),
                  "mcb_batch_add")
        self.shapes[cid.value] = (params.nz, params.ny, params.nx)
        return cid.value

    def add_case(self, case, **overrides):
        p = case.params
        for k, v in overrides.items():
            setattr(p, k, v)
        return self.add(p, case_bubbles(case), case.width)

    def remove(self, cid):
        self._chk(lib().mcb_batch_remove(self._h, cid), "mcb_batch_remove")
        self.shapes.pop(cid, None)

    def step(self, n=1):
        self._chk(lib().mcb_batch_step(self._h, int(n)), "mcb_batch_step")

    def step_timed(self, n):
        ms = ctypes.c_double()
        self._chk(lib().mcb_batch_step_timed(self._h, int(n), ctypes.byref(ms)), "mcb_batch_step_timed")
        return ms.value

    def compute_macros(self, cid):
        self._chk(lib().mcb_batch_compute_macros(self._h, cid), "mcb_batch_compute_macros")

    def read(self, cid, name):
        shape = self.shapes[cid]
        fid = FIELDS[name]
        cells = shape[0] * shape[1] * shape[2]
        n = 19 * cells if fid >= 16 else cells
        out = np.empty(n, dtype=np.float32)
        self._chk(lib().mcb_batch_read_field(self._h, cid, fid, _fptr(out), n), f"read {name}")
        return out.reshape(((19,) if fid >= 16 else ()) + shape)

    def write(self, cid, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32).ravel()
        self._chk(lib().mcb_batch_write_field(self._h, cid, FIELDS[name], _fptr(a), a.size), f"write {name}")

    def observables(self, cid):
        self.compute_macros(cid)
        return {k: self.read(cid, k) for k in ("phi", "rho", "ux", "uy", "uz")}

    def reduce(self, cid):
        s, k = ctypes.c_double(), ctypes.c_double()
        st = lib().mcb_batch_reduce(self._h, cid, ctypes.byref(s), ctypes.byref(k))
        if st not in (OK, ERR_NONFINITE):
            self._chk(st, "mcb_batch_reduce")
        return s.value, k.value, st == OK

    def save_state(self, cid):
        """State image of one case (mcb_batch_save_state)."""
        n = ctypes.c_size_t()
        self._chk(lib().mcb_batch_state_size(self._h, cid, ctypes.byref(n)), "mcb_batch_state_size")
        buf = ctypes.create_string_buffer(n.value)
        self._chk(lib().mcb_batch_save_state(self._h, cid, buf, n.value), "mcb_batch_save_state")
        return buf.raw

    def load_state(self, cid, image):
        buf = _image_buffer(image)
        self._chk(lib().mcb_batch_load_state(self._h, cid, buf, len(buf)), "mcb_batch_load_state")

    def add_state(self, image):
        """A new case resuming from a state image; returns its id."""
        buf = _image_buffer(image)
        cid = ctypes.c_int()
        self._chk(lib().mcb_batch_add_state(self._h, buf, len(buf), ctypes.byref(cid)), "mcb_batch_add_state")
        p, _ = state_info(image)
        self.shapes[cid.value] = (p.nz, p.ny, p.nx)
        return cid.value

    def case_steps(self, cid):
        v = ctypes.c_longlong()
        self._chk(lib().mcb_batch_case_steps(self._h, cid, ctypes.byref(v)), "mcb_batch_case_steps")
        return v.value

    def set_tiling(self, cid, tx, ty, zc):
        self._chk(lib().mcb_batch_set_tiling(self._h, cid, tx, ty, zc), "mcb_batch_set_tiling")

    def synchronize(self):
        self._chk(lib().mcb_batch_synchronize(self._h), "mcb_batch_synchronize")

    def __len__(self):
        return lib().mcb_batch_count(self._h)

    def nbytes(self):
        return lib().mcb_batch_bytes(self._h)


def serve_trace(n, rate_per_s, step_scale=1.0, seed=0):
    """Requests of the serving workload with Poisson arrivals (mcb_serve_trace)."""
    reqs = (Request * n)()
    check(lib().mcb_serve_trace(ctypes.c_uint64(seed), n, rate_per_s, step_scale, reqs), what="mcb_serve_trace")
    return reqs


def serve(reqs, *, storage=STORE_FP16, sequential=False, budget=0, max_cases=0, tick_steps=0, device=0):
    """Replay a request trace (mcb_serve_run); returns (results array, stats)."""
    cfg = ServeConfig()
    cfg.device, cfg.storage, cfg.hbm_budget = device, storage, budget
    cfg.max_cases, cfg.tick_steps, cfg.sequential = max_cases, tick_steps, int(sequential)
    n = len(reqs)
    res = (RequestResult * n)()
    st = ServeStats()
    check(lib().mcb_serve_run(ctypes.byref(cfg), reqs, n, res, ctypes.byref(st)), what="mcb_serve_run")
    return res, st


def tilings():
    out = []
    for i in range(lib().mcb_tiling_count()):
        a, b = ctypes.c_int(), ctypes.c_int()
        check(lib().mcb_tiling_at(i, ctypes.byref(a), ctypes.byref(b)))
        out.append((a.value, b.value))
    return out


def op_consts(params):
    c = OpConsts()
    check(lib().mcb_op_consts_from_params(ctypes.byref(params), ctypes.byref(c)), what="op consts")
    return c


class DeviceArray:
    """A device buffer owned by Python (for driving mcb_ops.h directly)."""

    def __init__(self, nbytes, device=0):
        p = ctypes.c_void_p()
        check(lib().mcb_device_alloc(device, nbytes, ctypes.byref(p)), what="device alloc")
        self.ptr, self.nbytes = p, nbytes

    @classmethod
    def from_numpy(cls, a, device=0):
        a = np.ascontiguousarray(a, dtype=np.float32)
        d = cls(a.nbytes, device)
        check(lib().mcb_memcpy_htod(d.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes), what="htod")
        return d

    def to_numpy(self, shape):
        out = np.empty(shape, dtype=np.float32)
        assert out.nbytes == self.nbytes
        check(lib().mcb_memcpy_dtoh(out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes), what="dtoh")
        return out

    def free(self):
        if self.ptr:
            lib().mcb_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def op(name, *args):
    """Call mcb_<name> of mcb_ops.h: device arrays, then NX, NY, NZ, consts, stream."""
    fn = getattr(lib(), "mcb_" + name)
    npt = OP_POINTERS[name]
    ptrs = [a.ptr if isinstance(a, DeviceArray) else a for a in args[:npt]]
    nx, ny, nz, consts = args[npt:npt + 4]
    stream = args[npt + 4] if len(args) > npt + 4 else None
    check(fn(*ptrs, nx, ny, nz, ctypes.byref(consts), stream), what=f"mcb_{name}")
