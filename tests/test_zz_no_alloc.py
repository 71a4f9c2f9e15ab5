"""mcb_reduce_fields_noalloc against mcb_reduce, through the C ABI.

The allocation-free entry point runs mcb_reduce's kernels on caller-owned
device fields with its partial sums in the library's static device array.
Its two doubles must equal mcb_reduce's bit for bit on the same fields, on
two calls in a row, the second with a smaller lattice (the array keeps the
first call's partials, which the second must not see).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def fields_and_reduce(mcb, n, steps, seed):
    p = mcb.make_params(n + 3, n, n - 2, tau_f=0.6, tau_g=1.0, sigma=0.05)
    rng = np.random.default_rng(seed)
    c = (rng.uniform(0, n + 3), rng.uniform(0, n), rng.uniform(0, n - 2), n * 0.25)
    with mcb.Lattice(p) as lat:
        lat.init_bubbles([c], 3.0)
        lat.step(steps)
        lat.compute_macros()
        s, ke, ok = lat.reduce()
        assert ok
        f = {k: lat.read(k) for k in ("phi", "rho", "ux", "uy", "uz")}
    return p, f, (s, ke)


def noalloc_reduce(mcb, p, f):
    dev = {k: mcb.DeviceArray.from_numpy(v) for k, v in f.items()}
    out = mcb.DeviceArray(16)
    st = mcb.lib().mcb_reduce_fields_noalloc(dev["phi"].ptr, dev["rho"].ptr, dev["ux"].ptr, dev["uy"].ptr,
                                             dev["uz"].ptr, p.nx, p.ny, p.nz, out.ptr, None)
    mcb.check(st, what="mcb_reduce_fields_noalloc")
    mcb.check(mcb.lib().mcb_device_synchronize(), what="synchronize")
    got = np.zeros(2, np.float64)
    mcb.check(mcb.lib().mcb_memcpy_dtoh(got.ctypes.data_as(ctypes.c_void_p), out.ptr, 16), what="dtoh")
    for d in list(dev.values()) + [out]:
        d.free()
    return got


def test_noalloc_reduce_equals_mcb_reduce_on_two_calls(gpu_lib):
    mcb = gpu_lib
    for n, steps, seed in ((40, 12, 1), (17, 5, 2)):   # the second lattice is smaller
        p, f, want = fields_and_reduce(mcb, n, steps, seed)
        got = noalloc_reduce(mcb, p, f)
        assert got[0] == want[0] and got[1] == want[1], f"n={n}: {got} vs mcb_reduce {want}"


def test_noalloc_reduce_rejects_bad_sizes_without_launching(gpu_lib):
    mcb = gpu_lib
    d = mcb.DeviceArray(64)
    L = mcb.lib()
    # 2^30 cells or more: rejected before any pointer is read.
    assert L.mcb_reduce_fields_noalloc(d.ptr, d.ptr, d.ptr, d.ptr, d.ptr, 1024, 1024, 1024, d.ptr, None) == \
        mcb.ERR_INVALID_ARG
    assert L.mcb_reduce_fields_noalloc(d.ptr, d.ptr, d.ptr, d.ptr, d.ptr, 0, 4, 4, d.ptr, None) == mcb.ERR_INVALID_ARG
    assert L.mcb_reduce_fields_noalloc(None, d.ptr, d.ptr, d.ptr, d.ptr, 2, 2, 2, d.ptr, None) == mcb.ERR_INVALID_ARG
    d.free()
